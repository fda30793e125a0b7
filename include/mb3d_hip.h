/* mb3d_hip.h -- C ABI of the MI355X (gfx950) HIP backend of mb3d-rs's
 * per-pixel distance-estimated (DE) fractal ray marcher.
 *
 * The backend is a drop-in sibling of the reference's wgpu backend
 * (`src/gpu.rs`): it takes the renderer's derived march parameters
 * (`CalcParams`, flattened into `mb3d_hip_params`) and the post job
 * (`PostJob`: hard shadows, DE ambient occlusion) and returns the G-buffer
 * (the 20-byte card record of `src/gpu.rs:3452-3463`) and/or the raw
 * buffers: f32 Z-buffer, u16 iteration counts, f32 normals.
 *
 * Conventions follow the reference's C ABI (capi/include/mb3d.h:1-16,
 * capi/src/lib.rs:18-41): plain structs and pointers, an int status, and a
 * thread-local text for the last failure (mb3d_hip_last_error).  No C++
 * exception crosses this boundary.  Pointers are not kept after a call
 * returns.  One device per process (mb3d_hip_init); calls from several
 * threads are serialised inside the library.
 */
#ifndef MB3D_HIP_H
#define MB3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MB3D_HIP_ABI_VERSION 3
#define MB3D_HIP_MAX_SLOTS 6
#define MB3D_HIP_MAX_LIGHTS 6

/* ---- status codes (the reference's Option/Result of gpu::march,
 *      src/gpu.rs:2296-2404, as integers) ---- */
#define MB3D_HIP_OK 0
#define MB3D_HIP_CANCELLED 1      /* the cancel callback asked to stop ("cancelled") */
#define MB3D_HIP_UNSUPPORTED 2    /* the scene needs a feature this backend lacks: run the CPU path */
#define MB3D_HIP_OVERFLOW 3       /* the reference's verdict for > 1/500 of the pixels overflowing single
                                     precision (src/gpu.rs:2350-2375).  Never returned here: a ray whose FP32
                                     DE is not finite goes on in the f64 finish, which decides as the CPU does
                                     (counted in mb3d_hip_counters.overflow) */
#define MB3D_HIP_ERR_INVALID (-1) /* bad argument */
#define MB3D_HIP_ERR_DEVICE (-2)  /* HIP runtime error, no device */
#define MB3D_HIP_ERR_NOT_INIT (-3)

/* ---- formula kinds of a hybrid slot: the built-in formulas
 *      (`Formula`, src/formulas.rs:28-68) ---- */
#define MB3D_HIP_FORMULA_NONE 0
#define MB3D_HIP_FORMULA_INTPOW 1        /* 'Integer Power' (triplex sine bulb), powers 2..8 */
#define MB3D_HIP_FORMULA_AMAZINGBOX 2    /* 'Amazing Box' (Mandelbox) */
#define MB3D_HIP_FORMULA_REALPOWER 3     /* 'Real Power' */
#define MB3D_HIP_FORMULA_QUATERNION 4    /* 'Quaternion' (4D; DE option 4) */
#define MB3D_HIP_FORMULA_TRICORN 5       /* 'Tricorn' */
#define MB3D_HIP_FORMULA_BULBOX 6        /* 'Bulbox' (power-2 bulb / Amazing Box blend) */
#define MB3D_HIP_FORMULA_FOLDINGINTPOW 7 /* 'Folding Int Pow' */
#define MB3D_HIP_FORMULA_AEXIONC 8       /* 'Aexion C' */
#define MB3D_HIP_FORMULA_OPTIONS 8       /* values in mb3d_hip_formula.options */

/* ---- DE modes ---- */
#define MB3D_HIP_DE_REFERENCE 0   /* the reference's: numeric gradient (option 0) or analytic box (11) */
#define MB3D_HIP_DE_ANALYTIC_P8 1 /* extension: running derivative, 0.5 r ln r / dr (one IntPow slot only) */

/* ---- iterate precision of the march ---- */
#define MB3D_HIP_PREC_AUTO 0 /* FP32 for an analytic DE unless the zoom is beyond it (the reference's
                              * needs_double rule); FP64 for a numeric DE */
#define MB3D_HIP_PREC_F32 1  /* FP32 iterate; FP64 ray position, stop decision, bisection, normals.
                              * Integer Power 8 and Amazing Box (and their hybrids) only: the other
                              * formulas, other powers and the 4D hybrids march in FP64 */
#define MB3D_HIP_PREC_F64 2  /* FP64 everywhere */

/* One slot of the alternating hybrid (`FormulaEntry` + `Formula`,
 * src/scene.rs, src/formulas.rs:28-68). */
typedef struct mb3d_hip_formula {
    int32_t kind;       /* MB3D_HIP_FORMULA_* */
    int32_t iterations; /* consecutive iterations of this slot (iItCount); < 0: not counted */
    int32_t power;      /* IntPow: the integer power */
    int32_t reserved;
    double z_mul;       /* IntPow: Z multiplier */
    double scale;       /* AmazingBox: scale */
    double min_r;       /* AmazingBox: Min R */
    double fold;        /* AmazingBox: fold */
    /* the other kinds: the values of Formula::options() in its order
     * (src/formulas.rs:308-357), integers and switches as 0/1 doubles:
     *   Real Power     float power, Z multiplier
     *   Quaternion     YW multiplier, W add
     *   Tricorn        Z multiplier, CZ multiplier
     *   Bulbox         box scale, box Min R, box fold, bulb scaling, R threshold, R threshold 2
     *   Folding Int Pow  integer power, Z multiplier, R fold
     *   Aexion C       float power, Z multiplier, rotate C, conditional phi, float power C,
     *                  Cz multiplier, power C on |Z - C|, mode (0..31) */
    double options[MB3D_HIP_FORMULA_OPTIONS];
} mb3d_hip_formula;

/* The subset of the reference's `Scene` (src/scene.rs:398-532) that feeds
 * `CalcParams::new` for the built-in formulas, plus two declared
 * extensions (`de_mode`, `step_cap`).  Defaults: mb3d_hip_scene_default. */
typedef struct mb3d_hip_scene {
    int32_t abi_version;
    int32_t width, height;
    int32_t iterations;        /* max iterations */
    int32_t min_iterations;
    int32_t smooth_normals;    /* 0..8 */
    int32_t bin_search_steps;  /* bStepsafterDEStop */
    int32_t first_step_random; /* 0/1 */
    int32_t step_sub_de_stop;  /* 0/1 */
    int32_t vary_de_stop_on_fov; /* 0/1 */
    int32_t normals_on_de;     /* 0/1 */
    int32_t color_option;      /* 0..5 */
    int32_t dfog_on_it;
    int32_t repeat_from;
    int32_t n_formulas;        /* 1..6 */
    int32_t de_mode;           /* MB3D_HIP_DE_* */
    int32_t step_cap;          /* extension: max march steps per ray (0: the reference's guard of 2e6) */
    int32_t julia;             /* 0/1: Julia mode, the constant is julia_c (Scene::julia) */
    int32_t interpolation;     /* 0/1: the interpolation hybrid of the first two formulas
                                  (Scene::interpolation, src/scene.rs:510) */
    float interpolation_w[2];  /* its weights (normalised to a sum of 1 by the set-up) */
    /* the DE combination of two hybrid parts (Scene::decomb, DeCombParams,
     * src/scene.rs:2178-2200): 0 none, 1 min, 2 max, 3 max with the second
     * part inverted, 4 smooth min (linear), 5 smooth min */
    int32_t decomb;
    int32_t decomb_end1, decomb_start2, decomb_end2, decomb_repeat2; /* slot indices from 0 */
    int32_t decomb_iterations2; /* the second part's maximum iterations */
    float decomb_smooth;        /* the smooth minima's distance (scene units) */
    int32_t part2_placed;       /* 0/1: the second part placed by part2_* (Scene::part2_place) */
    int32_t color_on_it;        /* the colour from the orbit after color_on_it - 1 iterations
                                   (1: the start point; 0: the last DE's orbit), Scene::color_on_it */
    double rstop;              /* bailout radius (not squared); <= 0: the formulas' default */
    double zoom;
    double mid[3];
    double z_start, z_end;
    double vgrads[3][3];       /* camera matrix rows: screen x, screen y, view direction */
    double fov_y;              /* degrees */
    double de_stop;
    double z_step_div;         /* raystep */
    double raystep_limiter;
    double color_mul;
    double julia_c[4];         /* the Julia constant (x, y, z, w); w is also J4 of the 3D hybrids */
    double rot_4d[3];          /* XW, YW, ZW rotations of the 4D hybrids (radians) */
    double part2_offset[3];    /* the second part's middle (Placement, src/scene.rs:49-56) */
    double part2_rotation[3];  /* its rotation about x, y, z (degrees) */
    double part2_scale;        /* its size */
    mb3d_hip_formula formulas[MB3D_HIP_MAX_SLOTS];
} mb3d_hip_scene;

/* The derived march parameters: the reference's `CalcParams`
 * (src/calc.rs:21-135) for the built-in formulas, with its f32-rounded
 * fields kept as f32.  Filled by mb3d_hip_params_from_scene (a
 * restatement of CalcParams::new, src/calc.rs:375-898) or by the caller
 * from its own CalcParams. */
typedef struct mb3d_hip_params {
    int32_t abi_version;
    int32_t width, height;
    int32_t n_slots;
    int32_t slot_kind[MB3D_HIP_MAX_SLOTS];
    int32_t slot_iters[MB3D_HIP_MAX_SLOTS];
    int32_t slot_uncounted[MB3D_HIP_MAX_SLOTS];
    int32_t slot_ade[MB3D_HIP_MAX_SLOTS];
    int32_t slot_power[MB3D_HIP_MAX_SLOTS];
    int32_t de_option;       /* 0 numeric, 11 analytic box, 32 analytic p8 (extension) */
    int32_t is_custom_de;
    int32_t end_to, repeat_from;
    int32_t max_it, min_it;
    int32_t sm_normals, de_add_steps;
    int32_t first_step_random, normals_on_de;
    int32_t color_option, dfog_on_it;
    int32_t step_cap;
    int32_t four_d;          /* 1: a 4D hybrid (DE options 4..6, src/calc.rs:475) */
    int32_t do_julia;        /* Julia mode: the constant is julia_c */
    int32_t ipol;            /* 1: the interpolation hybrid of slots 0 and 1 (HybridMode::Ipol3D/4D) */
    /* the DE combination (DeComb, src/calc.rs:224-236): its kind (0 none)
     * and the second part (HybridPart): slots start2..end2 repeating from
     * repeat2, its iteration limit, DE option, analytic flag, 4D mode */
    int32_t decomb;
    int32_t start2, end2, repeat2, max_it2;
    int32_t de_option2, is_custom_de2, four_d2;
    int32_t placed;          /* the second part at place_* (Place, src/calc.rs:238-253) */
    int32_t color_on_it;     /* CalcParams::color_on_it (do_color_on_it, src/calc.rs:2010-2025) */
    double slot_z_mul[MB3D_HIP_MAX_SLOTS];
    double slot_scale[MB3D_HIP_MAX_SLOTS];
    double slot_min_r[MB3D_HIP_MAX_SLOTS];
    double slot_fold[MB3D_HIP_MAX_SLOTS];
    float d_de_scale, de_stop, z_step_div, mh04zsd;
    float de_stop_factor, ln_rstop, de_offset, de_offset006;
    float mcts_m, ms_de_sub, d_col_plus, col_var_de_stop_mul;
    float fovx_off, fovx_mul, pl_optic_z, mct_color_mul;
    float fhln[MB3D_HIP_MAX_SLOTS];
    float ipol_w[2];         /* the interpolation weights, normalised (CalcParams::ipol, src/calc.rs:25) */
    float d_de_scale2;       /* the second part's DE scale */
    float reserved2;
    double d_rstop, rstop3d, fov_y, zcorr, zc_mul, zend, step_width;
    double vgrads[3][3];
    double ystart[3];
    /* the options of the kinds other than Integer Power and Amazing Box
     * (mb3d_hip_formula.options) */
    double slot_opt[MB3D_HIP_MAX_SLOTS][MB3D_HIP_FORMULA_OPTIONS];
    double smatrix4[4][4];   /* build_smatrix4 of rot_4d, f32-rounded (src/math.rs:178-204) */
    double julia_c[4];       /* `ju` (src/calc.rs:52) */
    double decomb_smooth;    /* the smooth minima's distance in steps */
    double place_offset[3], place_m[3][3], place_scale; /* m (p - offset); distances times scale */
} mb3d_hip_params;

/* The post passes that run with the march (`PostJob`, src/calc.rs:187-205;
 * `ShadowJob`, `HsLight` src/calc.rs:2705-2732; `DeaoParams`
 * src/deao.rs:14-29).  Directional lights only. */
typedef struct mb3d_hip_post_job {
    int32_t n_lights;                      /* 0: no shadow pass */
    int32_t light_idx[MB3D_HIP_MAX_LIGHTS]; /* header light index: bit 0x400 << idx of `shadow` */
    double light_vec[MB3D_HIP_MAX_LIGHTS][3]; /* HSvecs: world direction towards the light, length -step_width */
    float shadow_max_len_mul;
    int32_t deao;            /* 0: no DE ambient occlusion */
    int32_t deao_quality;    /* 0..3 = 3, 7, 17, 33 rays (src/deao.rs:75-105); 4 = the 8-ray set (extension) */
    float deao_max_len;
    int32_t deao_first_step_random; /* the first step of each ray times 0.5..2 drawn from the row's generator,
                                       in the CPU's order (src/deao.rs:262-270, 348-351) */
    int32_t deao_dither;     /* 0: off, 1: 2x2, 2: 3x3 dithering of the ray directions (src/deao.rs:318-342);
                                not with quality 4 */
} mb3d_hip_post_job;

/* Output buffers of rows [row0, row0 + rows) of the picture, each row
 * `width` pixels.  Any pointer may be NULL (that output is skipped).  With
 * `on_device` = 0 they are host memory, else device memory of the
 * library's device. */
typedef struct mb3d_hip_out {
    float *depth;       /* march depth mzz in step widths (x step_width = scene units along the ray);
                           +inf: background; 0: the ray starts inside the set */
    double *depth64;    /* the same in f64 */
    uint16_t *iters;    /* it_result at the surface point (after the bisection); 0 background */
    float *normals;     /* 3 per pixel: unit normal in the view (vgrads) frame; 0 background */
    uint32_t *gbuf;     /* 5 words per pixel: the card record of src/gpu.rs:3452-3463 */
    int32_t on_device;
    int32_t reserved;
} mb3d_hip_out;

typedef struct mb3d_hip_mlp mb3d_hip_mlp;

/* Launch and precision options of one call (zeros: the defaults). */
typedef struct mb3d_hip_options {
    int32_t precision;     /* MB3D_HIP_PREC_* */
    int32_t tile_w, tile_h; /* pixel tile of the work queue (default 8 x 8; powers of two up to 64) */
    int32_t waves_per_cu;  /* persistent grid: waves per CU (default 16) */
    int32_t it_split;      /* extension: FP32 march steps after which a ray goes on in the f64 finish (0: never) */
    float guard_band;      /* FP32 stop decisions within (1 + guard_band) x DE stop go to FP64 (0: 0.02) */
    /* the learned warm start (extension; NULL: the plain march).  A ray
     * whose hit probability is at least mlp_conf_hit (0: 0.5; above 1: no
     * hit jumps) starts at the predicted hit distance less mlp_margin_rel of
     * it (0: 0.15) and mlp_margin_abs scene units.  The start is kept when
     * the exact DE there clears the stop distance's hand-off guard and a
     * tenth of the predicted distance left, with the iteration limit not
     * reached; otherwise the ray marches from its start plane.  A ray whose
     * hit probability is at most mlp_conf_miss (0: no miss jumps) is
     * background without a march -- unverified, so opt-in. */
    const mb3d_hip_mlp *mlp;
    float mlp_conf_hit, mlp_conf_miss, mlp_margin_rel, mlp_margin_abs;
    /* extension: idle lanes a wave of the FP32 march gathers before it
     * refills them with new work items (0: the default, 1: every turn) */
    int32_t refill_min;
    /* extension, with the warm start: a verified miss decision.  Each tile
     * of the queue marches one coarse cone ray (the cone march,
     * src/cone.rs:1-31, src/calc.rs:2357-2435) whose reach covers the
     * tile's rays; its depth is a safe start for the tile's rays, and a
     * cone that passes the end of the view makes the tile background
     * (0: off) */
    int32_t mlp_cone;
    /* extension, with the warm start and the cone rays: the network as the
     * cone pass's gate.  The network runs once per tile, on the tile's
     * middle ray (the cone ray's), ahead of the cone pass; a tile whose hit
     * probability is at least mlp_cone_gate (in (0, 1); 0: off) marches no
     * cone ray and its rays march plainly -- as for a tile whose cone ray
     * meets a surface -- so the picture is the one without the gate.  The
     * gain: the cone rays of tiles that hold a surface, where they decide
     * nothing.  With hit jumps off (mlp_conf_hit above 1) the picture is
     * the plain march's. */
    float mlp_cone_gate;
    int32_t reserved_o;
} mb3d_hip_options;

/* ---- painting: the G-buffer lit and coloured (src/lighting.rs; the
 *      reference paints on the CPU only, src/gpu.rs:6) ---- */
#define MB3D_HIP_PAINT_U8 0  /* 8 bit per channel: paint (src/render.rs:2499) */
#define MB3D_HIP_PAINT_U16 1 /* 16 bit: paint16 (src/render.rs:2512) */
#define MB3D_HIP_PAINT_F32 2 /* float, 1 = white, not clipped above white: paint_f32 (src/render.rs:2518) */

/* One light of the lighting record (`Light`, src/lighting.rs:95-123). */
typedef struct mb3d_hip_light {
    int32_t on;
    int32_t spec_func;          /* specular exponent 2 << spec_func (0..7) */
    int32_t diff_func;          /* diffuse function 0..3 */
    int32_t relative_to_object; /* angles relative to the object (rotated with the view) */
    int32_t hs_enabled;         /* hard shadows enabled for this light */
    int32_t positional;         /* `position` instead of the angles */
    int32_t visible;            /* visible light source: 0 none, 2, 4 (old style), 6, 8.. (glow) */
    int32_t map;                /* light map number: 0 (light maps are not supported) */
    uint8_t color[3];
    uint8_t reserved;
    float amplitude;
    double x_angle, y_angle;    /* radians */
    double position[3];         /* scene coordinates of a positional light */
} mb3d_hip_light;

/* A surface colour of the palette (`PaletteColor`, src/lighting.rs:149-154). */
typedef struct mb3d_hip_palette_color {
    uint16_t position;
    uint8_t diffuse[3];
    uint8_t specular[3];
} mb3d_hip_palette_color;

/* The lighting record (`Lighting`, src/lighting.rs:159-288) with the
 * scene's shadow state applied (effective_lighting, src/render.rs:2275-2283:
 * hs_calced, hs_soft, hs_set_cos).  Defaults: mb3d_hip_lighting_default.
 * Light maps, background pictures, the diffuse colour map and the
 * background as ambient light are not supported (set-up refuses them). */
typedef struct mb3d_hip_lighting {
    mb3d_hip_light lights[MB3D_HIP_MAX_LIGHTS];
    mb3d_hip_palette_color palette[10];
    uint16_t interior_pos[4];   /* the 4 interior colours: positions, */
    uint8_t interior_col[4][3]; /* colours */
    uint8_t interior_spec[4];   /* and specular amounts */
    uint8_t amb_top[3], amb_bottom[3];
    uint8_t depth_col[3], depth_col2[3];   /* depth colour at the top / bottom of the picture */
    uint8_t dyn_fog_col[3], dyn_fog_col2[3];
    uint8_t palette_alpha[10];  /* used when palette_alpha_set (else the brightest specular component) */
    uint8_t fine_col_adj[2];    /* used when fine_col_adj_on: 0..60, 30 neutral */
    uint8_t reserved_b[2];
    float fog_offset, fog_far_offset, depth_fog, diffuse, dyn_fog, specular, ambient;
    float color_start, color_end, amb_shadow, ind_light, gamma, roughness, var_col_z;
    float interior_start, interior_end, diffuse_shadowing;
    int32_t color_cycling, color_on_otrap, far_fog, depth_func, fine_col_adj_on;
    int32_t hs_calced;          /* bit i: light i's hard shadow is in the G-buffer */
    int32_t hs_soft, hs_set_cos, internal_gamma2, no_col_ipol, amb_rel_obj;
    int32_t dfog_options, ex_mode, palette_alpha_set;
    int32_t vol_light;          /* -1: none; else the light whose volumetric light the shadow words hold
                                   (Scene::vol_light, src/scene.rs:514) */
    int32_t has_amb_frame;      /* amb_frame / amb_fov (Lighting::amb_frame) */
    int32_t bg_image, diff_map, bg_ambient; /* unsupported features: 0 */
    int32_t reserved;
    double amb_frame[3][3];
    double amb_fov;
} mb3d_hip_lighting;

/* One derived light (`LightVal`, src/lighting.rs:582-603). */
typedef struct mb3d_hip_light_val {
    int32_t idx, on, sub_amb_sh, hs_calced;
    int32_t hs_mask;            /* 0x400 << idx, or 0xFFFF for the soft shadow */
    int32_t pow_func, diff_func, positional, visible;
    float ln[3];                /* view-space direction, or position relative to the scene middle */
    float col[3];
    float lmax_l, pos_z, pos_lp;
} mb3d_hip_light_val;

/* The painter's derived light values (`LightVals`, src/lighting.rs:605-698)
 * for the supported features, lights in paint order (SortLights).  Filled
 * by mb3d_hip_paint_setup, or by the caller from its own LightVals. */
typedef struct mb3d_hip_light_vals {
    int32_t n_lights;
    int32_t col_cycling, col_on_otrap, far_fog, depth_func, gamma_h, sqr, no_col_ipol;
    int32_t amb_rel_obj, has_amb_frame, dfog_options, ex_mode, vol_light, has_ground;
    int32_t col_pos[10];
    int32_t icol_pos[4];
    mb3d_hip_light_val lights[MB3D_HIP_MAX_LIGHTS];
    float amb_col[3], amb_col2[3], depth_col[3], depth_col2[3], dyn_fog_col[3], dyn_fog_col2[3];
    float s_depth, s_shad, s_shad_gr, s_shad_zmul, s_dyn_fog_mul, s_amb_shad, s_diff, s_spec;
    float s_ind_light_reflect, s_col_zmul, s_c_start, s_c_mul, s_ci_start, s_ci_mul, s_roughness_factor;
    float s_diffuse_shadowing, s_gamma;
    float col_dif[10][3], col_spe[10][3], col_spe_a[10], s_c_div[10];
    float col_int[4][3], s_ic_div[4], col_int_spec[4];
    float amb_frame[3][3], amb_fov;
    float ground_dif[3], ground_spe[3]; /* the ground plane's colours (LightVals::set_ground) */
} mb3d_hip_light_vals;

/* The painter's camera (`PaintCamera`, src/lighting.rs:726-745). */
typedef struct mb3d_hip_paint_camera {
    int32_t width, height;
    int32_t planar;             /* optic: 0 common, 1 rectilinear, 2 panorama, 3 fisheye */
    int32_t reserved;
    float fov, aspect, x_off, pl_optic_z;
    double zcorr, zc_mul, step_width, zz_stmit_dif;
    float m[3][3];              /* unit view matrix (rows: picture x, y, view direction) */
    float start_pos[3], x_add[3], y_add[3];
} mb3d_hip_paint_camera;

/* Work counters of the last call on this thread: the algorithmic work that
 * the roofline figures are computed from (each lane counts its own). */
typedef struct mb3d_hip_counters {
    uint64_t pixels;
    uint64_t hits;
    uint64_t steps;          /* march steps */
    uint64_t de_evals_f32;   /* DE evaluations with the FP32 iterate */
    uint64_t de_evals_f64;   /* DE evaluations with the FP64 iterate */
    uint64_t its_p8_f32, its_box_f32;  /* formula iterations by kind and precision: Integer Power, */
    uint64_t its_p8_f64, its_box_f64;  /* Amazing Box, */
    uint64_t its_other_f64;            /* the other formulas (FP64 only) */
    uint64_t its_normals_f32;          /* of its_p8_f32 + its_box_f32: the normals pass's FP32 stencil */
    uint64_t overflow;       /* pixels whose FP32 values overflowed */
    uint64_t jumps;          /* warm starts taken */
    uint64_t rollbacks;      /* warm starts rejected by the exact DE (overshoot) */
    double kernel_ms;        /* device time of the march kernels (HIP events) */
    double post_ms;          /* device time of the post kernels */
    double march_f32_ms;     /* of kernel_ms: the FP32 march (k_march32) */
    double finish_f64_ms;    /* of kernel_ms: the FP64 finish and the normals */
    double paint_ms;         /* device time of the last mb3d_hip_paint's kernel (HIP events) */
    uint64_t cone_rays;      /* cone rays marched (one per tile, mlp_cone) */
    uint64_t cone_background; /* pixels made background by their tile's cone ray */
    double mlp_ms;           /* of kernel_ms: the warm start's network (k_mlp_propose) */
    double cone_ms;          /* of kernel_ms: the cone rays (k_cone) */
    uint64_t mlp_rays;       /* rays the network ran on (tiles the cone rays found empty skip it) */
    uint64_t its_cone_f64;   /* of the FP64 iterations: the cone rays' (k_cone) */
    uint64_t de_cone_f64;    /* of de_evals_f64: the cone rays' */
    double normals_ms;       /* of finish_f64_ms: the normals pass (k_normals) */
    uint64_t steps_f64;      /* of steps: the f64 finish's (k_finish64) */
} mb3d_hip_counters;

typedef void (*mb3d_hip_progress_fn)(void *user, int64_t done, int64_t total);
typedef int (*mb3d_hip_cancel_fn)(void *user); /* non-zero: stop */

/* Why the last failing call on this thread failed (capi/src/lib.rs:18-26).
 * Valid until the next failing call on the thread; not to be freed. */
const char *mb3d_hip_last_error(void);

/* The ABI version the library was built with (MB3D_HIP_ABI_VERSION). */
int32_t mb3d_hip_abi_version(void);

/* Opens HIP device `device` (the reference's one global device, OnceLock,
 * src/gpu.rs:1042-1088).  Idempotent for the same device. */
int32_t mb3d_hip_init(int32_t device);

/* Releases the library's device buffers. */
void mb3d_hip_shutdown(void);

/* Scene defaults (Scene::default, src/scene.rs:626-700) with one Integer
 * Power p8 slot (src/formulas.rs:266-306). */
void mb3d_hip_scene_default(mb3d_hip_scene *scene);

/* Formula::default_for (src/formulas.rs:266-306): the defaults of the
 * built-in formula `name` ("Integer Power", "Real Power", "Quaternion",
 * "Tricorn", "Amazing Box", "Bulbox", "Folding Int Pow", "Aexion C", or
 * the reference's aliases), one counted iteration.  MB3D_HIP_UNSUPPORTED
 * for other names (the .m3f formulas). */
int32_t mb3d_hip_formula_default(const char *name, mb3d_hip_formula *formula);

/* CalcParams::new (src/calc.rs:375-898) for the built-in formulas. */
int32_t mb3d_hip_params_from_scene(const mb3d_hip_scene *scene, mb3d_hip_params *params);

/* gpu::unsupported (src/gpu.rs:445-477): 0 when this backend can march
 * `params`, else MB3D_HIP_UNSUPPORTED with the reason in `reason`. */
int32_t mb3d_hip_supported(const mb3d_hip_params *params, char *reason, size_t reason_len);

/* pixel_seed (src/calc.rs:1125-1140): the random state of pixel (x, y). */
int32_t mb3d_hip_pixel_seed(int64_t y, int64_t x);

/* gpu::march / gpu::march_bands (src/gpu.rs:2296, 1586): marches rows
 * [row0, row0 + rows) of the picture, runs the post job on them and fills
 * `out`.  `progress` and `cancel` (may be NULL) are called on the calling
 * thread between kernel slices. */
int32_t mb3d_hip_march(const mb3d_hip_params *params, const mb3d_hip_post_job *post,
                       int32_t row0, int32_t rows, const mb3d_hip_out *out,
                       const mb3d_hip_options *options, mb3d_hip_progress_fn progress,
                       mb3d_hip_cancel_fn cancel, void *user);

/* No counterpart in the reference: mb3d_hip_march for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It marches rows [row0, row0 +
 * rows) into `out`, which must be device memory (out->on_device = 1), with
 * no post job and no network (options->mlp NULL); its lists, tables and
 * counters live in a 256 MiB scratch array of the library's own device
 * memory, held per device until the process ends.  A call needs about
 * 180 bytes a pixel and rejects more than the array holds (about 1.3 M
 * pixels at 256 CUs: a 1080p picture marches in two bands of rows) with
 * MB3D_HIP_ERR_INVALID.  Calls that overlap on one device race on the
 * array. */
int32_t mb3d_hip_march_noalloc(const mb3d_hip_params *params, int32_t row0, int32_t rows,
                               const mb3d_hip_out *out, const mb3d_hip_options *options);

/* gpu::post_on_card (src/gpu.rs:1649): the shadow and DE-AO passes on a
 * G-buffer of rows [row0, row0 + rows) marched elsewhere (5 words per
 * pixel, in and out). */
int32_t mb3d_hip_post(const mb3d_hip_params *params, const mb3d_hip_post_job *post,
                      int32_t row0, int32_t rows, uint32_t *gbuf, int32_t on_device,
                      const mb3d_hip_options *options);

/* The rays of the pixels of rows [row0, row0 + rows), as the warm-start
 * network takes them (host memory, row-major): `rays` 6 floats per pixel
 * (start point on the start plane, unit direction), `step_len` the length
 * of one march unit along the ray (a depth in march units times it is the
 * distance in scene units).  Either output may be NULL.  The start point
 * and direction are the march's (calc_vgrads_fov, calc_start_pos,
 * src/calc.rs:1561-1580 and 1673-1689). */
int32_t mb3d_hip_rays(const mb3d_hip_params *params, int32_t row0, int32_t rows, float *rays, float *step_len);

/* calc::de_at (src/calc.rs:2695-2703) at `n` points (3 doubles each, host
 * memory): the f64 DE in world units and the iteration count there.  Either
 * output may be NULL. */
int32_t mb3d_hip_de_at(const mb3d_hip_params *params, const double *points, int64_t n, double *de,
                       int32_t *its);

/* ---- batches (no reference counterpart in one call: the reference renders
 * animation frames one by one, a card each, src/frames.rs:467-520) ----
 * Up to 64 frames of any sizes with the same formula and DE mode, marched
 * in one persistent launch of each pass.  The outputs are device buffers
 * holding the frames one after another (frame f starts at the sum of the
 * earlier frames' width x height pixels). */
int32_t mb3d_hip_march_batch(const mb3d_hip_params *frames, int32_t n, const mb3d_hip_out *out,
                             const mb3d_hip_options *options);

/* A batch captured once into a HIP graph and replayed: create, launch any
 * number of times (each launch waits for its last replay), destroy. */
typedef struct mb3d_hip_plan mb3d_hip_plan;
int32_t mb3d_hip_plan_create(const mb3d_hip_params *frames, int32_t n, const mb3d_hip_out *out,
                             const mb3d_hip_options *options, mb3d_hip_plan **plan);
int32_t mb3d_hip_plan_launch(mb3d_hip_plan *plan, int32_t repeats);
void mb3d_hip_plan_destroy(mb3d_hip_plan *plan);

/* The autotuner: sweeps the tile shape (8x8 .. 32x8), the persistent grid
 * (waves per CU), the iteration-cap split (it_split 0, 128, 256) and the
 * refill threshold (refill_min 8, 16) on graphs of the batch and returns
 * the fastest options in `best` (its other fields are kept) and their
 * kernel time. */
int32_t mb3d_hip_autotune(const mb3d_hip_params *frames, int32_t n, const mb3d_hip_out *out,
                          mb3d_hip_options *best, double *best_ms);

/* ---- the warm-start network (no counterpart in the reference) ----
 * An MLP proposing each ray's hit distance from its origin and direction
 * (sin/cos encoding, 3 x 64 ReLU, outputs distance and hit logit), trained
 * on this backend's own HIP forward and backward kernels from (ray, exact
 * distance) pairs of the plain march, run as BF16 MFMA inside the march.
 * Rays: n x 6 floats (origin, unit direction); targets: n x 2 floats
 * (distance / 4, hit 0/1). */
#define MB3D_HIP_MLP_PARAMS 13634
int32_t mb3d_hip_mlp_create(uint64_t seed, mb3d_hip_mlp **net);
void mb3d_hip_mlp_destroy(mb3d_hip_mlp *net);
/* copies the f32 parameters out (get) and/or in (set); n = MB3D_HIP_MLP_PARAMS */
int32_t mb3d_hip_mlp_params(mb3d_hip_mlp *net, float *get, const float *set, int64_t n);
/* mean loss and its gradient over the batch (one forward and backward) */
int32_t mb3d_hip_mlp_grad(mb3d_hip_mlp *net, const float *rays, const float *targets, int64_t n, float *loss,
                          float *grad);
/* Adam training on the pairs: `steps` minibatches of `batch` (in order,
 * wrapping), learning rate `lr` with cosine decay; the loss of each step
 * into loss_curve (may be NULL) */
int32_t mb3d_hip_mlp_train(mb3d_hip_mlp *net, const float *rays, const float *targets, int64_t n, int32_t steps,
                           int32_t batch, float lr, float *loss_curve);
/* the network's outputs (n x 2: distance / 4, hit logit) in f32 or with the
 * march's BF16 MFMA path */
int32_t mb3d_hip_mlp_infer(mb3d_hip_mlp *net, const float *rays, int64_t n, int32_t bf16_mfma, float *out);

/* The work counters of the last march/post call on this thread (and the
 * last paint's kernel time). */
int32_t mb3d_hip_counters_get(mb3d_hip_counters *out);

/* ---- secondary rays (the reference's reflection batches: gpu::reflect_batch,
 *      src/gpu.rs:2521-2737, on the CPU SrTracer::march_ray,
 *      src/reflect.rs:1464-1674) ---- */

/* A ray from a surface point (ReflectTask, src/gpu.rs:2426-2441). */
typedef struct mb3d_hip_ray_task {
    double origin[3];   /* the point the ray leaves */
    float dir[3];       /* its direction, one step width long */
    float zz;           /* the view depth there (steps): the DE stop grows with it */
    uint32_t pixel;     /* its pixel (y * width + x): the pixel's view ray projects the ray on the depth */
    int32_t inside;     /* 1: it starts inside the set (no march: open air) */
    float max_len;      /* the longest march (steps); 0: to the far end (zend) */
    int32_t ao_q;       /* the AO quality at the hit, 0..3 (deao_1pos; the bounces take lower ones) */
} mb3d_hip_ray_task;

/* The shadow and occlusion pass at the rays' hits (ReflectOpts,
 * src/gpu.rs:2459-2481; calc_hs and deao_1pos, src/reflect.rs:805-1090). */
typedef struct mb3d_hip_ray_opts {
    int32_t n_lights;                        /* the lights with hard shadows (0: none) */
    int32_t light_idx[MB3D_HIP_MAX_LIGHTS];  /* header light index: the hit's shadow bit 0x400 << idx */
    int32_t light_positional[MB3D_HIP_MAX_LIGHTS];
    int32_t light_visible[MB3D_HIP_MAX_LIGHTS];
    float light_lmax[MB3D_HIP_MAX_LIGHTS];   /* lmax_l: how far a positional light reaches */
    double light_vec[MB3D_HIP_MAX_LIGHTS][3]; /* HSvec of a global light (world frame, -step_width long
                                                towards the light), the scene position of a positional one */
    float soft_radius;  /* > 0: one soft shadow of the last light, of this radius */
    float hs_mul;       /* the shadow rays' length multiplier */
    int32_t ao;         /* 1: DE ambient occlusion at the hit, of each task's ao_q */
    float ao_max_len;
    int32_t ao_first_random; /* 0 (a random first AO step is not supported) */
    int32_t reserved;
} mb3d_hip_ray_opts;

#define MB3D_HIP_RAY_HIT 1          /* the ray met a surface */
#define MB3D_HIP_RAY_DE_LIMITED 2   /* the DE stopped it (not the iteration limit) */

/* What a ray met (ReflectHit, src/gpu.rs:2447-2457; RayHit,
 * src/reflect.rs:324-336). */
typedef struct mb3d_hip_ray_hit {
    uint32_t gbuf[5];   /* the hit's card record: normal, depth code of the ray length, step count, colour,
                           AO 5000 (none); open air: background with the step count */
    int32_t flags;      /* MB3D_HIP_RAY_* */
    double pos[3];      /* where the ray ended */
    double zz;          /* the view depth there */
    double zz2;         /* the ray length after the surface search */
    double zz2_paint;   /* the ray length where the march stopped */
    float step_count;
    int32_t iters;      /* the iteration count at the hit */
} mb3d_hip_ray_hit;

/* Marches `n` rays, then runs the shadow and occlusion pass of `opts` at
 * their hits (NULL: none; AO 5000 and no shadow bits, as the reference
 * without them).  depth_sqrt: the hit's depth code from
 * sqrt((zz2 + zz)^2 - zz), as the painter's far fog or internal gamma 2 ask
 * (src/reflect.rs:1634-1638), else from zz2.  The AO is returned as the card
 * returns it: the caller scales it by the bounce (amb * 8 / (8 + bounce),
 * src/reflect.rs:2270-2275).  on_device: `tasks` and `hits` are device
 * memory, else host memory. */
int32_t mb3d_hip_march_rays(const mb3d_hip_params *params, const mb3d_hip_ray_task *tasks, int64_t n,
                            int32_t depth_sqrt, const mb3d_hip_ray_opts *opts, mb3d_hip_ray_hit *hits,
                            int32_t on_device);

/* ---- the Monte Carlo renderer's marches (gpu::mc_batch,
 *      src/gpu.rs:3013-3190; on the CPU McMarch::run, src/mc.rs:478-977) ---- */

#define MB3D_HIP_MC_RAY 0      /* a ray of CalcRay, from pos along dir, at most len */
#define MB3D_HIP_MC_SHADOW 1   /* a shadow ray of CalcHSMC towards a light, len long */
#define MB3D_HIP_MC_THROUGH 2  /* the way through a transparent object, len left */

#define MB3D_HIP_MC_INSIDE 1     /* task and result: inside rendering (the DE from inside the set) */
#define MB3D_HIP_MC_PROBE 2      /* task: a primary ray, which first looks whether it starts inside */
#define MB3D_HIP_MC_TRANS 4      /* task: a shadow ray finds, colours and returns the surface it hits */
#define MB3D_HIP_MC_NO_GROUND 8  /* task: the ground plane is left out (there is none on this backend) */
#define MB3D_HIP_MC_DE_LIMITED 2 /* result: the DE stopped the ray (not the iteration limit) */

/* A march of the tracer (McReq, src/mc.rs:410-438). */
typedef struct mb3d_hip_mc_task {
    double pos[3];      /* the start */
    double dir[3];      /* the direction, one step width long */
    double zz;          /* the view depth at the start (steps): the DE stop grows with it */
    double zz2;         /* the ray's length so far (a shadow ray: the view depth along it) */
    double zzplus;      /* the view depth per step */
    double len;         /* the longest march, the shadow ray's length or the length left */
    double vfov[3];     /* the sample's view vector (the volumetric light's, which is not on this backend) */
    int32_t kind;       /* MB3D_HIP_MC_RAY, _SHADOW or _THROUGH */
    int32_t seed;       /* the random seed (GetRand) */
    uint32_t flags;     /* MB3D_HIP_MC_INSIDE | _PROBE | _TRANS | _NO_GROUND */
    float ms_de_stop;   /* the DE stop of a THROUGH task */
} mb3d_hip_mc_task;

/* Where a march ended (McRes, src/mc.rs:440-468). */
typedef struct mb3d_hip_mc_res {
    double pos[3];      /* where it ended */
    double zz;          /* the view depth there (RAY) */
    double zz2;         /* the ray length where the march stopped (RAY), the depth along it (SHADOW) */
    double len;         /* the length left (SHADOW, THROUGH) */
    double normals[3];  /* the surface's unit normal */
    float dyn_fog;      /* the dynamic fog's step count (RAY) */
    float ms_de_stop;   /* the DE stop where it ended (SHADOW) */
    int32_t kind;       /* 0 open air (or a shadow ray without a surface), 1 a surface, 2 a primary ray
                           that starts inside the set */
    int32_t seed;       /* the random seed after the march */
    uint16_t si_grad, otrap;  /* the surface's gradient and orbit trap colour */
    uint32_t flags;     /* MB3D_HIP_MC_INSIDE | MB3D_HIP_MC_DE_LIMITED (dIFS formulas are not on this backend:
                           no dIFS flags) */
} mb3d_hip_mc_res;

/* Marches `n` tasks of the Monte Carlo renderer.  secant: the secant search
 * refines a ray's surface (else the bisection).  The card's own batch
 * returns f32 positions and depths unless the scene needs FP64; this one
 * returns every result in f64.  A primary ray that starts inside the set
 * ends there (kind 2): the turn-around of inside-and-outside rendering
 * (src/mc.rs:692-697) is off, as this backend's scenes have no such
 * option.  on_device: `tasks` and `res` are device memory, else host
 * memory. */
int32_t mb3d_hip_mc_batch(const mb3d_hip_params *params, const mb3d_hip_mc_task *tasks, int64_t n, int32_t secant,
                          mb3d_hip_mc_res *res, int32_t on_device);

/* ---- the G-buffer's file records ---- */

/* gbuffer::pack / unpack (src/gbuffer.rs:60-98): n card records (5 u32
 * each, as mb3d_hip_out.gbuf) to and from n 18-byte little-endian records
 * -- normal[3] i16, zpos_fine u32, shadow, amb_shadow, si_gradient, otrap
 * u16 -- the pixel block of an .m3i (write_m3i, src/m3p.rs:1846-1853).
 * The card record's overflow flag (bit 16 of the last word) is not kept.
 * on_device: both buffers are device memory of the library's device, else
 * host memory. */
#define MB3D_HIP_REC_BYTES 18
int32_t mb3d_hip_gbuf_pack(const uint32_t *gbuf, int64_t n, uint8_t *records, int32_t on_device);
int32_t mb3d_hip_gbuf_unpack(const uint8_t *records, int64_t n, uint32_t *gbuf, int32_t on_device);

/* ---- painting ---- */

/* Lighting::default (src/lighting.rs:290-392): MB3D's start-up light
 * preset; vol_light -1. */
void mb3d_hip_lighting_default(mb3d_hip_lighting *lighting);

/* The painter's camera and light values of a scene (paint_camera and
 * light_vals, src/render.rs:2340-2380 and 2285-2319, with LightVals::new,
 * set_vol_light, rotate_object_lights and place_lights,
 * src/lighting.rs:963-1320), for the common optic without stereo.
 * MB3D_HIP_UNSUPPORTED (reason in mb3d_hip_last_error) for light maps,
 * background pictures, the diffuse map or the background as ambient. */
int32_t mb3d_hip_paint_setup(const mb3d_hip_scene *scene, const mb3d_hip_params *params,
                             const mb3d_hip_lighting *lighting, mb3d_hip_paint_camera *camera,
                             mb3d_hip_light_vals *light_vals);

/* The colours of the pixels of rectangle (x0, y0, w, h) of the picture
 * from their G-buffer records (5 words per pixel, the card record of
 * src/gpu.rs:3452-3463, row-major, w per row): LightVals::pixel_color_s
 * (src/lighting.rs:1739-1760, shade 1800-2150) for every pixel, as
 * paint_core's loop (src/render.rs:2718-2736).  `rgb` holds 3 channels per
 * pixel in `format` (MB3D_HIP_PAINT_*).  on_device: both buffers are
 * device memory of the library's device (the G-buffer as the march left
 * it), else host memory. */
int32_t mb3d_hip_paint(const mb3d_hip_paint_camera *camera, const mb3d_hip_light_vals *light_vals,
                       const uint32_t *gbuf, int32_t x0, int32_t y0, int32_t w, int32_t h,
                       int32_t format, void *rgb, int32_t on_device);

#ifdef __cplusplus
}
#endif

#endif /* MB3D_HIP_H */
