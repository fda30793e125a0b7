"""ctypes mirror of include/mb3d_hip.h and a thin Python front of the C ABI.

The product is the C-ABI library ``libmb3d_hip.so`` built next to this file
(HIP kernels for gfx950 plus the C++ host side).  This module only declares
the structs, loads the library (it fails loudly when it is missing) and
offers the scene builders of the measured workloads.

Scene fields mirror the reference's ``Scene`` (src/scene.rs:398-532) and its
defaults (src/scene.rs:626-700); camera matrices follow
``vgrads_from_angles`` (src/math.rs:102-123).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmb3d_hip.so")

ABI_VERSION = 3
MAX_SLOTS = 6
MAX_LIGHTS = 6

OK, CANCELLED, UNSUPPORTED, OVERFLOW = 0, 1, 2, 3
ERR_INVALID, ERR_DEVICE, ERR_NOT_INIT = -1, -2, -3
FORMULA_NONE, FORMULA_INTPOW, FORMULA_AMAZINGBOX = 0, 1, 2
FORMULA_REALPOWER, FORMULA_QUATERNION, FORMULA_TRICORN = 3, 4, 5
FORMULA_BULBOX, FORMULA_FOLDINGINTPOW, FORMULA_AEXIONC = 6, 7, 8
FORMULA_OPTIONS = 8
# Formula::all_names (src/formulas.rs:249-262) and the kinds they map to
FORMULA_NAMES = {
    "Integer Power": FORMULA_INTPOW, "Real Power": FORMULA_REALPOWER, "Quaternion": FORMULA_QUATERNION,
    "Tricorn": FORMULA_TRICORN, "Amazing Box": FORMULA_AMAZINGBOX, "Bulbox": FORMULA_BULBOX,
    "Folding Int Pow": FORMULA_FOLDINGINTPOW, "Aexion C": FORMULA_AEXIONC,
}
DE_REFERENCE, DE_ANALYTIC_P8 = 0, 1
PREC_AUTO, PREC_F32, PREC_F64 = 0, 1, 2


class Formula(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("iterations", C.c_int32),
        ("power", C.c_int32),
        ("reserved", C.c_int32),
        ("z_mul", C.c_double),
        ("scale", C.c_double),
        ("min_r", C.c_double),
        ("fold", C.c_double),
        ("options", C.c_double * FORMULA_OPTIONS),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("iterations", C.c_int32),
        ("min_iterations", C.c_int32),
        ("smooth_normals", C.c_int32),
        ("bin_search_steps", C.c_int32),
        ("first_step_random", C.c_int32),
        ("step_sub_de_stop", C.c_int32),
        ("vary_de_stop_on_fov", C.c_int32),
        ("normals_on_de", C.c_int32),
        ("color_option", C.c_int32),
        ("dfog_on_it", C.c_int32),
        ("repeat_from", C.c_int32),
        ("n_formulas", C.c_int32),
        ("de_mode", C.c_int32),
        ("step_cap", C.c_int32),
        ("julia", C.c_int32),
        ("interpolation", C.c_int32),
        ("interpolation_w", C.c_float * 2),
        ("decomb", C.c_int32),
        ("decomb_end1", C.c_int32),
        ("decomb_start2", C.c_int32),
        ("decomb_end2", C.c_int32),
        ("decomb_repeat2", C.c_int32),
        ("decomb_iterations2", C.c_int32),
        ("decomb_smooth", C.c_float),
        ("part2_placed", C.c_int32),
        ("color_on_it", C.c_int32),
        ("rstop", C.c_double),
        ("zoom", C.c_double),
        ("mid", C.c_double * 3),
        ("z_start", C.c_double),
        ("z_end", C.c_double),
        ("vgrads", (C.c_double * 3) * 3),
        ("fov_y", C.c_double),
        ("de_stop", C.c_double),
        ("z_step_div", C.c_double),
        ("raystep_limiter", C.c_double),
        ("color_mul", C.c_double),
        ("julia_c", C.c_double * 4),
        ("rot_4d", C.c_double * 3),
        ("part2_offset", C.c_double * 3),
        ("part2_rotation", C.c_double * 3),
        ("part2_scale", C.c_double),
        ("formulas", Formula * MAX_SLOTS),
    ]


S6i = C.c_int32 * MAX_SLOTS
S6d = C.c_double * MAX_SLOTS


class Params(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("n_slots", C.c_int32),
        ("slot_kind", S6i),
        ("slot_iters", S6i),
        ("slot_uncounted", S6i),
        ("slot_ade", S6i),
        ("slot_power", S6i),
        ("de_option", C.c_int32),
        ("is_custom_de", C.c_int32),
        ("end_to", C.c_int32),
        ("repeat_from", C.c_int32),
        ("max_it", C.c_int32),
        ("min_it", C.c_int32),
        ("sm_normals", C.c_int32),
        ("de_add_steps", C.c_int32),
        ("first_step_random", C.c_int32),
        ("normals_on_de", C.c_int32),
        ("color_option", C.c_int32),
        ("dfog_on_it", C.c_int32),
        ("step_cap", C.c_int32),
        ("four_d", C.c_int32),
        ("do_julia", C.c_int32),
        ("ipol", C.c_int32),
        ("decomb", C.c_int32),
        ("start2", C.c_int32),
        ("end2", C.c_int32),
        ("repeat2", C.c_int32),
        ("max_it2", C.c_int32),
        ("de_option2", C.c_int32),
        ("is_custom_de2", C.c_int32),
        ("four_d2", C.c_int32),
        ("placed", C.c_int32),
        ("color_on_it", C.c_int32),
        ("slot_z_mul", S6d),
        ("slot_scale", S6d),
        ("slot_min_r", S6d),
        ("slot_fold", S6d),
        ("d_de_scale", C.c_float),
        ("de_stop", C.c_float),
        ("z_step_div", C.c_float),
        ("mh04zsd", C.c_float),
        ("de_stop_factor", C.c_float),
        ("ln_rstop", C.c_float),
        ("de_offset", C.c_float),
        ("de_offset006", C.c_float),
        ("mcts_m", C.c_float),
        ("ms_de_sub", C.c_float),
        ("d_col_plus", C.c_float),
        ("col_var_de_stop_mul", C.c_float),
        ("fovx_off", C.c_float),
        ("fovx_mul", C.c_float),
        ("pl_optic_z", C.c_float),
        ("mct_color_mul", C.c_float),
        ("fhln", C.c_float * MAX_SLOTS),
        ("ipol_w", C.c_float * 2),
        ("d_de_scale2", C.c_float),
        ("reserved2", C.c_float),
        ("d_rstop", C.c_double),
        ("rstop3d", C.c_double),
        ("fov_y", C.c_double),
        ("zcorr", C.c_double),
        ("zc_mul", C.c_double),
        ("zend", C.c_double),
        ("step_width", C.c_double),
        ("vgrads", (C.c_double * 3) * 3),
        ("ystart", C.c_double * 3),
        ("slot_opt", (C.c_double * FORMULA_OPTIONS) * MAX_SLOTS),
        ("smatrix4", (C.c_double * 4) * 4),
        ("julia_c", C.c_double * 4),
        ("decomb_smooth", C.c_double),
        ("place_offset", C.c_double * 3),
        ("place_m", (C.c_double * 3) * 3),
        ("place_scale", C.c_double),
    ]


class PostJob(C.Structure):
    _fields_ = [
        ("n_lights", C.c_int32),
        ("light_idx", C.c_int32 * MAX_LIGHTS),
        ("light_vec", (C.c_double * 3) * MAX_LIGHTS),
        ("shadow_max_len_mul", C.c_float),
        ("deao", C.c_int32),
        ("deao_quality", C.c_int32),
        ("deao_max_len", C.c_float),
        ("deao_first_step_random", C.c_int32),
        ("deao_dither", C.c_int32),
    ]


class Out(C.Structure):
    _fields_ = [
        ("depth", C.c_void_p),
        ("depth64", C.c_void_p),
        ("iters", C.c_void_p),
        ("normals", C.c_void_p),
        ("gbuf", C.c_void_p),
        ("on_device", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("tile_w", C.c_int32),
        ("tile_h", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("it_split", C.c_int32),
        ("guard_band", C.c_float),
        ("mlp", C.c_void_p),
        ("mlp_conf_hit", C.c_float),
        ("mlp_conf_miss", C.c_float),
        ("mlp_margin_rel", C.c_float),
        ("mlp_margin_abs", C.c_float),
        ("refill_min", C.c_int32),
        ("mlp_cone", C.c_int32),
        ("mlp_cone_gate", C.c_float),
        ("reserved_o", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("pixels", C.c_uint64),
        ("hits", C.c_uint64),
        ("steps", C.c_uint64),
        ("de_evals_f32", C.c_uint64),
        ("de_evals_f64", C.c_uint64),
        ("its_p8_f32", C.c_uint64),
        ("its_box_f32", C.c_uint64),
        ("its_p8_f64", C.c_uint64),
        ("its_box_f64", C.c_uint64),
        ("its_other_f64", C.c_uint64),
        ("its_normals_f32", C.c_uint64),
        ("overflow", C.c_uint64),
        ("jumps", C.c_uint64),
        ("rollbacks", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("post_ms", C.c_double),
        ("march_f32_ms", C.c_double),
        ("finish_f64_ms", C.c_double),
        ("paint_ms", C.c_double),
        ("cone_rays", C.c_uint64),
        ("cone_background", C.c_uint64),
        ("mlp_ms", C.c_double),
        ("cone_ms", C.c_double),
        ("mlp_rays", C.c_uint64),
        ("its_cone_f64", C.c_uint64),
        ("de_cone_f64", C.c_uint64),
        ("normals_ms", C.c_double),
        ("steps_f64", C.c_uint64),
    ]


RAY_HIT, RAY_DE_LIMITED = 1, 2


class RayTask(C.Structure):
    _fields_ = [
        ("origin", C.c_double * 3),
        ("dir", C.c_float * 3),
        ("zz", C.c_float),
        ("pixel", C.c_uint32),
        ("inside", C.c_int32),
        ("max_len", C.c_float),
        ("ao_q", C.c_int32),
    ]


class RayOpts(C.Structure):
    _fields_ = [
        ("n_lights", C.c_int32),
        ("light_idx", C.c_int32 * MAX_LIGHTS),
        ("light_positional", C.c_int32 * MAX_LIGHTS),
        ("light_visible", C.c_int32 * MAX_LIGHTS),
        ("light_lmax", C.c_float * MAX_LIGHTS),
        ("light_vec", (C.c_double * 3) * MAX_LIGHTS),
        ("soft_radius", C.c_float),
        ("hs_mul", C.c_float),
        ("ao", C.c_int32),
        ("ao_max_len", C.c_float),
        ("ao_first_random", C.c_int32),
        ("reserved", C.c_int32),
    ]


class RayHit(C.Structure):
    _fields_ = [
        ("gbuf", C.c_uint32 * 5),
        ("flags", C.c_int32),
        ("pos", C.c_double * 3),
        ("zz", C.c_double),
        ("zz2", C.c_double),
        ("zz2_paint", C.c_double),
        ("step_count", C.c_float),
        ("iters", C.c_int32),
    ]


# numpy views of the two records
RAY_TASK_DTYPE = np.dtype([("origin", "<f8", 3), ("dir", "<f4", 3), ("zz", "<f4"), ("pixel", "<u4"),
                           ("inside", "<i4"), ("max_len", "<f4"), ("ao_q", "<i4")])
RAY_HIT_DTYPE = np.dtype([("gbuf", "<u4", 5), ("flags", "<i4"), ("pos", "<f8", 3), ("zz", "<f8"), ("zz2", "<f8"),
                          ("zz2_paint", "<f8"), ("step_count", "<f4"), ("iters", "<i4")])

# the Monte Carlo renderer's marches (McReq / McRes, src/mc.rs:410-468)
MC_RAY, MC_SHADOW, MC_THROUGH = 0, 1, 2
MC_INSIDE, MC_PROBE, MC_TRANS, MC_NO_GROUND = 1, 2, 4, 8
MC_DE_LIMITED = 2
MC_TASK_DTYPE = np.dtype([("pos", "<f8", 3), ("dir", "<f8", 3), ("zz", "<f8"), ("zz2", "<f8"), ("zzplus", "<f8"),
                          ("len", "<f8"), ("vfov", "<f8", 3), ("kind", "<i4"), ("seed", "<i4"), ("flags", "<u4"),
                          ("ms_de_stop", "<f4")])
MC_RES_DTYPE = np.dtype([("pos", "<f8", 3), ("zz", "<f8"), ("zz2", "<f8"), ("len", "<f8"), ("normals", "<f8", 3),
                         ("dyn_fog", "<f4"), ("ms_de_stop", "<f4"), ("kind", "<i4"), ("seed", "<i4"),
                         ("si_grad", "<u2"), ("otrap", "<u2"), ("flags", "<u4")])


PAINT_U8, PAINT_U16, PAINT_F32 = 0, 1, 2


class Light(C.Structure):
    _fields_ = [
        ("on", C.c_int32),
        ("spec_func", C.c_int32),
        ("diff_func", C.c_int32),
        ("relative_to_object", C.c_int32),
        ("hs_enabled", C.c_int32),
        ("positional", C.c_int32),
        ("visible", C.c_int32),
        ("map", C.c_int32),
        ("color", C.c_uint8 * 3),
        ("reserved", C.c_uint8),
        ("amplitude", C.c_float),
        ("x_angle", C.c_double),
        ("y_angle", C.c_double),
        ("position", C.c_double * 3),
    ]


class PaletteColor(C.Structure):
    _fields_ = [("position", C.c_uint16), ("diffuse", C.c_uint8 * 3), ("specular", C.c_uint8 * 3)]


_LIGHTING_FLOATS = ("fog_offset", "fog_far_offset", "depth_fog", "diffuse", "dyn_fog", "specular", "ambient",
                    "color_start", "color_end", "amb_shadow", "ind_light", "gamma", "roughness", "var_col_z",
                    "interior_start", "interior_end", "diffuse_shadowing")
_LIGHTING_INTS = ("color_cycling", "color_on_otrap", "far_fog", "depth_func", "fine_col_adj_on", "hs_calced",
                  "hs_soft", "hs_set_cos", "internal_gamma2", "no_col_ipol", "amb_rel_obj", "dfog_options",
                  "ex_mode", "palette_alpha_set", "vol_light", "has_amb_frame", "bg_image", "diff_map",
                  "bg_ambient", "reserved")


class Lighting(C.Structure):
    _fields_ = [
        ("lights", Light * MAX_LIGHTS),
        ("palette", PaletteColor * 10),
        ("interior_pos", C.c_uint16 * 4),
        ("interior_col", (C.c_uint8 * 3) * 4),
        ("interior_spec", C.c_uint8 * 4),
        ("amb_top", C.c_uint8 * 3),
        ("amb_bottom", C.c_uint8 * 3),
        ("depth_col", C.c_uint8 * 3),
        ("depth_col2", C.c_uint8 * 3),
        ("dyn_fog_col", C.c_uint8 * 3),
        ("dyn_fog_col2", C.c_uint8 * 3),
        ("palette_alpha", C.c_uint8 * 10),
        ("fine_col_adj", C.c_uint8 * 2),
        ("reserved_b", C.c_uint8 * 2),
    ] + [(n, C.c_float) for n in _LIGHTING_FLOATS] + [(n, C.c_int32) for n in _LIGHTING_INTS] + [
        ("amb_frame", (C.c_double * 3) * 3),
        ("amb_fov", C.c_double),
    ]


class LightVal(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("idx", "on", "sub_amb_sh", "hs_calced", "hs_mask", "pow_func",
                                         "diff_func", "positional", "visible")] + [
        ("ln", C.c_float * 3),
        ("col", C.c_float * 3),
        ("lmax_l", C.c_float),
        ("pos_z", C.c_float),
        ("pos_lp", C.c_float),
    ]


class LightVals(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("n_lights", "col_cycling", "col_on_otrap", "far_fog", "depth_func",
                                         "gamma_h", "sqr", "no_col_ipol", "amb_rel_obj", "has_amb_frame",
                                         "dfog_options", "ex_mode", "vol_light", "has_ground")] + [
        ("col_pos", C.c_int32 * 10),
        ("icol_pos", C.c_int32 * 4),
        ("lights", LightVal * MAX_LIGHTS),
    ] + [(n, C.c_float * 3) for n in ("amb_col", "amb_col2", "depth_col", "depth_col2", "dyn_fog_col",
                                      "dyn_fog_col2")] + [
        (n, C.c_float) for n in ("s_depth", "s_shad", "s_shad_gr", "s_shad_zmul", "s_dyn_fog_mul", "s_amb_shad",
                                 "s_diff", "s_spec", "s_ind_light_reflect", "s_col_zmul", "s_c_start", "s_c_mul",
                                 "s_ci_start", "s_ci_mul", "s_roughness_factor", "s_diffuse_shadowing", "s_gamma")
    ] + [
        ("col_dif", (C.c_float * 3) * 10),
        ("col_spe", (C.c_float * 3) * 10),
        ("col_spe_a", C.c_float * 10),
        ("s_c_div", C.c_float * 10),
        ("col_int", (C.c_float * 3) * 4),
        ("s_ic_div", C.c_float * 4),
        ("col_int_spec", C.c_float * 4),
        ("amb_frame", (C.c_float * 3) * 3),
        ("amb_fov", C.c_float),
        ("ground_dif", C.c_float * 3),
        ("ground_spe", C.c_float * 3),
    ]


class PaintCamera(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("planar", C.c_int32),
        ("reserved", C.c_int32),
        ("fov", C.c_float),
        ("aspect", C.c_float),
        ("x_off", C.c_float),
        ("pl_optic_z", C.c_float),
        ("zcorr", C.c_double),
        ("zc_mul", C.c_double),
        ("step_width", C.c_double),
        ("zz_stmit_dif", C.c_double),
        ("m", (C.c_float * 3) * 3),
        ("start_pos", C.c_float * 3),
        ("x_add", C.c_float * 3),
        ("y_add", C.c_float * 3),
    ]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_int64)
CANCEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)

# every symbol include/mb3d_hip.h declares, with its ctypes signature
SIGNATURES = {
    "mb3d_hip_last_error": (C.c_char_p, []),
    "mb3d_hip_abi_version": (C.c_int32, []),
    "mb3d_hip_init": (C.c_int32, [C.c_int32]),
    "mb3d_hip_shutdown": (None, []),
    "mb3d_hip_scene_default": (None, [C.POINTER(Scene)]),
    "mb3d_hip_formula_default": (C.c_int32, [C.c_char_p, C.POINTER(Formula)]),
    "mb3d_hip_params_from_scene": (C.c_int32, [C.POINTER(Scene), C.POINTER(Params)]),
    "mb3d_hip_supported": (C.c_int32, [C.POINTER(Params), C.c_char_p, C.c_size_t]),
    "mb3d_hip_pixel_seed": (C.c_int32, [C.c_int64, C.c_int64]),
    "mb3d_hip_march": (
        C.c_int32,
        [C.POINTER(Params), C.POINTER(PostJob), C.c_int32, C.c_int32, C.POINTER(Out),
         C.POINTER(Options), PROGRESS_FN, CANCEL_FN, C.c_void_p],
    ),
    "mb3d_hip_march_noalloc": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.POINTER(Out),
                                           C.POINTER(Options)]),
    "mb3d_hip_post": (
        C.c_int32,
        [C.POINTER(Params), C.POINTER(PostJob), C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
         C.POINTER(Options)],
    ),
    "mb3d_hip_counters_get": (C.c_int32, [C.POINTER(Counters)]),
    "mb3d_hip_de_at": (C.c_int32, [C.POINTER(Params), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mb3d_hip_rays": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "mb3d_hip_march_batch": (C.c_int32, [C.POINTER(Params), C.c_int32, C.POINTER(Out), C.POINTER(Options)]),
    "mb3d_hip_plan_create": (C.c_int32, [C.POINTER(Params), C.c_int32, C.POINTER(Out), C.POINTER(Options),
                                         C.POINTER(C.c_void_p)]),
    "mb3d_hip_plan_launch": (C.c_int32, [C.c_void_p, C.c_int32]),
    "mb3d_hip_plan_destroy": (None, [C.c_void_p]),
    "mb3d_hip_autotune": (C.c_int32, [C.POINTER(Params), C.c_int32, C.POINTER(Out), C.POINTER(Options),
                                      C.POINTER(C.c_double)]),
    "mb3d_hip_mlp_create": (C.c_int32, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "mb3d_hip_mlp_destroy": (None, [C.c_void_p]),
    "mb3d_hip_mlp_params": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "mb3d_hip_mlp_grad": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mb3d_hip_mlp_train": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                       C.c_float, C.c_void_p]),
    "mb3d_hip_mlp_infer": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
    "mb3d_hip_march_rays": (C.c_int32, [C.POINTER(Params), C.c_void_p, C.c_int64, C.c_int32, C.POINTER(RayOpts),
                                        C.c_void_p, C.c_int32]),
    "mb3d_hip_mc_batch": (C.c_int32, [C.POINTER(Params), C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int32]),
    "mb3d_hip_gbuf_pack": (C.c_int32, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]),
    "mb3d_hip_gbuf_unpack": (C.c_int32, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]),
    "mb3d_hip_lighting_default": (None, [C.POINTER(Lighting)]),
    "mb3d_hip_paint_setup": (C.c_int32, [C.POINTER(Scene), C.POINTER(Params), C.POINTER(Lighting),
                                         C.POINTER(PaintCamera), C.POINTER(LightVals)]),
    "mb3d_hip_paint": (C.c_int32, [C.POINTER(PaintCamera), C.POINTER(LightVals), C.c_void_p, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32]),
}

_lib = None


def load(path: str = LIB_PATH):
    """The product library; raises when it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libmb3d_hip.so is not built ({path}): run __graft_entry__.build()")
    # torch ships its own HIP runtime under another file name; loaded first,
    # it satisfies the library's libamdhip64.so.7 too, so that a process that
    # uses both (device buffers from torch) holds one HIP runtime, not two
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    e = load().mb3d_hip_last_error()
    return e.decode() if e else ""


def check(status: int, what: str) -> int:
    if status < 0 or status == UNSUPPORTED:
        raise RuntimeError(f"{what} failed ({status}): {last_error()}")
    return status


# ---------------------------------------------------------------------------
# scene helpers (harness side)


def vgrads_from_angles(ax: float, ay: float, az: float) -> list:
    """vgrads_from_angles (src/math.rs:102-123): rotate around X, then Y, then Z."""
    s1, c1 = math.sin(ax), math.cos(ax)
    s2, c2 = math.sin(ay), math.cos(ay)
    s3, c3 = math.sin(az), math.cos(az)
    m = []
    for i in range(3):
        x1, y1, z1 = (1.0, 0.0, 0.0) if i == 0 else (0.0, 1.0, 0.0) if i == 1 else (0.0, 0.0, 1.0)
        x2 = x1
        y2 = c1 * y1 + s1 * z1
        z2 = c1 * z1 - s1 * y1
        x1 = c2 * x2 + s2 * z2
        y1 = y2
        z1 = c2 * z2 - s2 * x2
        m.append([c3 * x1 + s3 * y1, c3 * y1 - s3 * x1, z1])
    return m


def scene_default() -> Scene:
    """Scene::default (src/scene.rs:626-700) with one Integer Power p8 slot."""
    s = Scene()
    s.abi_version = ABI_VERSION
    s.width, s.height = 800, 600
    s.iterations = 60
    s.min_iterations = 1
    s.smooth_normals = 0
    s.bin_search_steps = 8
    s.first_step_random = 1
    s.step_sub_de_stop = 0
    s.vary_de_stop_on_fov = 1
    s.normals_on_de = 1
    s.color_option = 0
    s.dfog_on_it = 0
    s.repeat_from = 0
    s.n_formulas = 1
    s.de_mode = DE_REFERENCE
    s.step_cap = 0
    s.rstop = 0.0
    s.zoom = 1.0
    s.mid[:] = [0.0, 0.0, 0.0]
    s.z_start, s.z_end = -2.0, 30.0
    for i in range(3):
        for j in range(3):
            s.vgrads[i][j] = 1.0 if i == j else 0.0
    s.fov_y = 30.0
    s.de_stop = 1.0
    s.z_step_div = 0.5
    s.raystep_limiter = 1.0
    s.color_mul = 1.0
    # DeCombParams::new (src/scene.rs:2196-2200), used when a combination is on
    s.decomb_end1, s.decomb_start2, s.decomb_end2, s.decomb_repeat2 = 0, 1, 5, 1
    s.decomb_iterations2 = s.iterations
    s.decomb_smooth = 0.5
    s.part2_scale = 1.0
    set_formulas(s, [intpow(8, 1)])
    return s


def intpow(power=8, iterations=1, z_mul=-1.0) -> Formula:
    f = Formula()
    f.kind, f.iterations, f.power, f.z_mul = FORMULA_INTPOW, iterations, power, z_mul
    return f


def amazing_box(scale=2.0, min_r=0.5, fold=1.0, iterations=1) -> Formula:
    f = Formula()
    f.kind, f.iterations, f.scale, f.min_r, f.fold = FORMULA_AMAZINGBOX, iterations, scale, min_r, fold
    return f


def formula(name: str, iterations=1, **options) -> Formula:
    """A built-in formula with its defaults (Formula::default_for,
    src/formulas.rs:266-306, through mb3d_hip_formula_default); keyword
    options override them by position name: for the kinds other than
    Integer Power and Amazing Box, ``options`` index names from OPTION_NAMES."""
    f = Formula()
    check(load().mb3d_hip_formula_default(name.encode(), C.byref(f)), f"formula {name!r}")
    f.iterations = iterations
    names = OPTION_NAMES.get(f.kind, ())
    for k, v in options.items():
        if k in names:
            f.options[names.index(k)] = v
        elif not names and k in ("power", "z_mul", "scale", "min_r", "fold"):
            setattr(f, k, v)
        else:
            raise KeyError(f"{name} has no option {k!r}")
    return f


# the option names of Formula::options (src/formulas.rs:308-357) for the
# kinds that keep their values in Formula.options
OPTION_NAMES = {
    FORMULA_REALPOWER: ("power", "z_mul"),
    FORMULA_QUATERNION: ("yw_mul", "w_add"),
    FORMULA_TRICORN: ("z_mul", "cz_mul"),
    FORMULA_BULBOX: ("scale", "min_r", "fold", "bulb_scale", "r_threshold", "r_threshold2"),
    FORMULA_FOLDINGINTPOW: ("power", "z_mul", "fold"),
    FORMULA_AEXIONC: ("power", "z_mul", "rot_c", "cond_phi", "power_c", "cz_mul", "powc_dist", "mode"),
}


def scene_preset(name: str, width=96, height=72) -> Scene:
    """Scene::preset (src/scene.rs:811-838): one built-in formula with its
    camera framing (z_start, zoom, iterations, pitch and yaw); the
    Quaternion preset is a Julia set at (-0.2, 0.6, 0.2, 0)."""
    s = scene_default()
    s.width, s.height = width, height
    f = formula(name)
    frames = {
        FORMULA_AMAZINGBOX: (-14.0, 0.18, 20, (0.35, -0.6)),
        FORMULA_BULBOX: (-14.0, 0.18, 20, (0.35, -0.6)),
        FORMULA_QUATERNION: (-3.0, 0.8, 12, (0.4, 0.3)),
        FORMULA_FOLDINGINTPOW: (-20.0, 0.15, 20, (0.5, 0.4)),
        FORMULA_TRICORN: (-3.0, 0.6, 12, (0.6, 0.2)),
        FORMULA_AEXIONC: (-3.0, 0.65, 10, (0.45, 0.3)),
    }
    zs, zoom, its, rot = frames.get(f.kind, (-2.5, 0.8, 12, (-0.55, 0.45)))
    if f.kind == FORMULA_QUATERNION:
        s.julia = 1
        s.julia_c[:] = [-0.2, 0.6, 0.2, 0.0]
    s.z_start, s.zoom, s.iterations = zs, zoom, its
    set_camera(s, vgrads_from_angles(rot[0], rot[1], 0.0))
    set_formulas(s, [f])
    return s


def set_formulas(s: Scene, fs: list) -> None:
    for i in range(MAX_SLOTS):
        s.formulas[i] = fs[i] if i < len(fs) else Formula()
    s.n_formulas = len(fs)


def set_camera(s: Scene, vg) -> None:
    for i in range(3):
        for j in range(3):
            s.vgrads[i][j] = vg[i][j]


def look_matrix(direction) -> list:
    """Camera rows (screen x, screen y, view) for a view direction."""
    d = np.asarray(direction, float)
    d = d / np.linalg.norm(d)
    up = np.array([0.0, 1.0, 0.0]) if abs(d[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    x = np.cross(up, d)
    x /= np.linalg.norm(x)
    y = np.cross(d, x)
    return [list(x), list(y), list(d)]


def scene_bulb(width, height, seed=1, de_mode=DE_ANALYTIC_P8, step_cap=500, bin_search=6) -> Scene:
    """The power-8 Mandelbulb workload: 60 iterations, bailout r^2 = 16,
    the camera's start plane 2.5 from the origin facing it (the reference's
    Integer Power preset framing, zoom 0.8, src/scene.rs:811-838), FOV 30,
    seeded yaw in [0, 2 pi) and pitch in [-0.5, 0.5]; no random first step;
    6 bisection steps; 500-step cap."""
    rng = np.random.default_rng(seed)
    yaw = float(rng.uniform(0.0, 2.0 * math.pi))
    pitch = float(rng.uniform(-0.5, 0.5))
    s = scene_default()
    s.width, s.height = width, height
    s.iterations = 60
    s.rstop = 4.0
    s.zoom = 0.8
    s.z_start, s.z_end = -2.5, 30.0
    s.fov_y = 30.0
    s.first_step_random = 0
    s.bin_search_steps = bin_search
    s.step_cap = step_cap
    s.de_mode = de_mode
    set_camera(s, vgrads_from_angles(pitch, yaw, 0.0))
    set_formulas(s, [intpow(8, 1)])
    return s


def scene_m3s_mandelbulb(width=800, height=600) -> Scene:
    """The reference's scene fixture tests/data/mandelbulb.m3s (the power-8
    Integer Power bulb close up, smooth normals 3), its keys transcribed:
    iterations 14, zoom 1.8, mid (0, -0.35, 0), z_start -2, z_end 30,
    rotation (-40, 30, 0) degrees, de_stop 0.8, raystep 0.4, smooth_normals
    3, bin_search 8, Integer Power 8 with z multiplier -1 and the reference's
    numeric DE.  (Its lights and colours are the painter's.)"""
    s = scene_default()
    s.width, s.height = width, height
    s.iterations = 14
    s.zoom = 1.8
    s.mid[:] = [0.0, -0.35, 0.0]
    s.z_start, s.z_end = -2.0, 30.0
    set_camera(s, vgrads_from_angles(math.radians(-40.0), math.radians(30.0), 0.0))
    s.de_stop = 0.8
    s.z_step_div = 0.4
    s.smooth_normals = 3
    s.bin_search_steps = 8
    s.de_mode = DE_REFERENCE
    set_formulas(s, [intpow(8, 1, -1.0)])
    return s


def scene_m3s_quaternion_julia(width=640, height=480) -> Scene:
    """The reference's scene fixture tests/data/quaternion_julia.m3s (the 4D
    Quaternion Julia set through the 4D hybrid loop with a 4D rotation), its
    keys transcribed: iterations 12, zoom 0.8, z_start -3, rotation (23, 17,
    0) degrees, rotation_4d (0, 20, 0) degrees, Julia at (-0.2, 0.6, 0.2, 0),
    Quaternion with yw multiplier -1 and w add 0.  (Its colours are the
    painter's.)"""
    s = scene_default()
    s.width, s.height = width, height
    s.iterations = 12
    s.zoom = 0.8
    s.z_start = -3.0
    set_camera(s, vgrads_from_angles(math.radians(23.0), math.radians(17.0), 0.0))
    s.rot_4d[:] = [0.0, math.radians(20.0), 0.0]
    s.julia = 1
    s.julia_c[:] = [-0.2, 0.6, 0.2, 0.0]
    set_formulas(s, [formula("Quaternion", yw_mul=-1.0, w_add=0.0)])
    return s


def scene_m3s_interpolation_boxes(width=640, height=480) -> Scene:
    """The reference's scene fixture tests/data/interpolation_boxes.m3s (two
    Amazing Boxes blended 60 : 40 in every iteration, MB3D's interpolation
    hybrid, analytic DE), its keys transcribed: iterations 30, zoom 0.35,
    z_start -8, rotation (25, -35, 0) degrees, smooth normals 2; Amazing Box
    (scale -1.5, min_r 0.5, fold 1) x2 and Amazing Box (scale 2.2, min_r 0.5,
    fold 1) x1."""
    s = scene_default()
    s.width, s.height = width, height
    s.iterations = 30
    s.interpolation = 1
    s.interpolation_w[:] = [0.6, 0.4]
    s.zoom = 0.35
    s.z_start = -8.0
    set_camera(s, vgrads_from_angles(math.radians(25.0), math.radians(-35.0), 0.0))
    s.smooth_normals = 2
    set_formulas(s, [amazing_box(-1.5, 0.5, 1.0, 2), amazing_box(2.2, 0.5, 1.0, 1)])
    return s


def _box_de(p, scale, min_r, fold, its):
    """Analytic Amazing Box DE |z|/|dr| at a point (harness: camera placement)."""
    z = np.array(p, float)
    c = z.copy()
    w = 1.0
    m2 = max(min_r, 1e-40) ** 2
    for _ in range(its):
        z = np.abs(z + fold) - (np.abs(z - fold) + z)
        r = float(z @ z)
        mul = scale / m2 if r < m2 else (scale if r > 1.0 else scale / r)
        w *= mul
        z = z * mul + c
        if float(z @ z) > 1024.0 ** 2:
            break
    return math.sqrt(float(z @ z)) / abs(w)


def scene_box(width, height, seed=1, step_cap=0) -> Scene:
    """The Amazing Box workload: scale -1.5, min R 0.5, fold 1, 30
    iterations, analytic DE; the camera's start plane centred at a seeded
    point inside the radius-4 sphere (re-drawn until it lies outside the set,
    its DE above 0.05) with a seeded view direction towards the inner half."""
    rng = np.random.default_rng(seed)
    for _ in range(1000):
        p = rng.uniform(-4.0, 4.0, 3)
        if p @ p > 16.0:
            continue
        if _box_de(p, -1.5, 0.5, 1.0, 30) < 0.05:
            continue
        break
    target = rng.uniform(-1.0, 1.0, 3)
    d = target - p
    s = scene_default()
    s.width, s.height = width, height
    s.iterations = 30
    s.zoom = 1.0
    s.mid[:] = list(p)
    s.z_start = float(p[2])
    s.z_end = s.z_start + 12.0
    s.first_step_random = 0
    s.step_cap = step_cap
    set_camera(s, look_matrix(d))
    set_formulas(s, [amazing_box(-1.5, 0.5, 1.0, 1)])
    return s


def scene_hybrid(width, height, seed=1) -> Scene:
    """The hybrid chain: Mandelbulb p8 and Amazing Box (scale 2) alternating
    1:2, 45 iterations, numeric DE (the reference's combined option 0)."""
    s = scene_bulb(width, height, seed=seed, de_mode=DE_REFERENCE, step_cap=500, bin_search=8)
    s.iterations = 45
    s.rstop = 0.0  # the formulas' default (1024 for the box)
    s.zoom = 0.5
    s.z_start = -4.0
    set_formulas(s, [intpow(8, 1), amazing_box(2.0, 0.5, 1.0, 2)])
    return s


def post_job(lights=(), deao_quality=None, shadow_max_len_mul=1.0, params: Params = None, deao_dither=0,
             deao_first_step_random=False) -> PostJob:
    """A post job with directional lights given as view-space directions
    towards the light (HSvecs = RotateVectorReverse(-StepWidth * LN, M),
    src/render.rs:2383-2417), and the DE AO of `DeaoParams`
    (src/deao.rs:14-29) when `deao_quality` is given."""
    j = PostJob()
    j.n_lights = len(lights)
    if lights:
        assert params is not None
        vg = np.array([[params.vgrads[i][k] for k in range(3)] for i in range(3)])
        m = vg / np.linalg.norm(vg, axis=1, keepdims=True)
        for i, ln in enumerate(lights):
            ln = np.asarray(ln, float)
            v = -ln / np.linalg.norm(ln) * params.step_width
            w = v @ m  # rotate_vector_reverse: sum_i v[i] * M[i]
            j.light_idx[i] = i
            for k in range(3):
                j.light_vec[i][k] = float(w[k])
    j.shadow_max_len_mul = shadow_max_len_mul
    if deao_quality is not None:
        j.deao = 1
        j.deao_quality = deao_quality
        j.deao_max_len = 1.0
        j.deao_dither = deao_dither
        j.deao_first_step_random = int(bool(deao_first_step_random))
    return j


def struct_to_dict(st) -> dict:
    """The fields of a ctypes struct of this ABI as plain Python values
    (nested arrays as lists), for storing outside the process."""
    def val(v):
        if isinstance(v, C.Array):
            return [val(x) for x in v]
        if isinstance(v, C.Structure):
            return struct_to_dict(v)
        return v
    return {name: val(getattr(st, name)) for name, _ in st._fields_}


def struct_from_dict(cls, d: dict):
    """A ctypes struct of type `cls` from struct_to_dict's output; fields the
    dict lacks stay zero, names it has that `cls` lacks are an error."""
    st = cls()
    names = {n for n, _ in cls._fields_}
    for k, v in d.items():
        if k not in names:
            raise KeyError(f"{cls.__name__} has no field {k!r}")
        cur = getattr(st, k)
        if isinstance(cur, C.Array):
            _fill_array(cur, v)
        elif isinstance(cur, C.Structure):
            setattr(st, k, struct_from_dict(type(cur), v))
        else:
            setattr(st, k, v)
    return st


def _fill_array(arr, vals):
    for i, v in enumerate(vals):
        if isinstance(arr[i], C.Array):
            _fill_array(arr[i], v)
        elif isinstance(arr[i], C.Structure):
            arr[i] = struct_from_dict(type(arr[i]), v)
        else:
            arr[i] = v


def params_from_scene(s: Scene) -> Params:
    p = Params()
    check(load().mb3d_hip_params_from_scene(C.byref(s), C.byref(p)), "params_from_scene")
    return p


def params_dict(p: Params) -> dict:
    out = {}
    for name, _ in Params._fields_:
        v = getattr(p, name)
        if hasattr(v, "__len__"):
            v = [list(x) if hasattr(x, "__len__") else x for x in v]
        out[name] = v
    return out


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def options(precision=PREC_AUTO, tile=(8, 8), waves_per_cu=16, guard_band=0.0, it_split=0, mlp=None,
            conf_hit=0.0, conf_miss=0.0, margin_rel=0.0, margin_abs=0.0, refill_min=0, cone=False,
            cone_gate=0.0) -> Options:
    """Options; `mlp` (an Mlp) turns the warm start on: hit jumps with the
    library's defaults (conf_hit 0: 0.5, above 1: off; margin_rel 0: 0.15),
    miss jumps (unverified) only with conf_miss > 0; `cone` adds the cone
    ray per tile (a verified safe start, and background for tiles it finds
    empty); `cone_gate` (with `mlp` and `cone`) skips the cone ray of a tile
    whose middle ray the network calls a hit with at least that
    probability.  `refill_min`: idle lanes a march wave gathers before
    refilling (0: the library's default)."""
    o = Options()
    o.precision = precision
    o.tile_w, o.tile_h = tile
    o.waves_per_cu = waves_per_cu
    o.guard_band = guard_band
    o.it_split = it_split
    o.refill_min = refill_min
    o.mlp_cone = 1 if cone else 0
    o.mlp_cone_gate = cone_gate
    if mlp is not None:
        o.mlp = mlp.h
        o.mlp_conf_hit, o.mlp_conf_miss = conf_hit, conf_miss
        o.mlp_margin_rel, o.mlp_margin_abs = margin_rel, margin_abs
    return o


def march(p: Params, post: PostJob = None, row0=0, rows=None, opts: Options = None, outputs=None):
    """mb3d_hip_march into new host arrays: dict of depth (f32), depth64,
    iters (u16), normals (n x 3 f32), gbuf (n x 5 u32) and the counters."""
    lib = load()
    rows = p.height - row0 if rows is None else rows
    n = rows * p.width
    want = outputs or ("depth", "depth64", "iters", "normals", "gbuf")
    res = {
        "depth": np.zeros(n, np.float32) if "depth" in want else None,
        "depth64": np.zeros(n, np.float64) if "depth64" in want else None,
        "iters": np.zeros(n, np.uint16) if "iters" in want else None,
        "normals": np.zeros((n, 3), np.float32) if "normals" in want else None,
        "gbuf": np.zeros((n, 5), np.uint32) if "gbuf" in want else None,
    }
    o = Out(*(_p(res[k]) for k in ("depth", "depth64", "iters", "normals", "gbuf")), 0, 0)
    st = lib.mb3d_hip_march(C.byref(p), C.byref(post) if post is not None else None, row0, rows, C.byref(o),
                            C.byref(opts) if opts is not None else None, PROGRESS_FN(0), CANCEL_FN(0), None)
    check(st, "mb3d_hip_march")
    res["status"] = st
    res["counters"] = counters()
    return res


def post(p: Params, job: PostJob, gbuf, row0=0, rows=None, opts: Options = None):
    lib = load()
    rows = p.height - row0 if rows is None else rows
    g = np.ascontiguousarray(gbuf, dtype=np.uint32).copy()
    st = lib.mb3d_hip_post(C.byref(p), C.byref(job), row0, rows, _p(g), 0,
                           C.byref(opts) if opts is not None else None)
    check(st, "mb3d_hip_post")
    return g


def march_rays(p: Params, tasks, depth_sqrt=False, opts: "RayOpts" = None):
    """mb3d_hip_march_rays on host records (a RAY_TASK_DTYPE array): the
    hits as a RAY_HIT_DTYPE array; `opts` the hits' shadow and AO pass."""
    t = np.ascontiguousarray(tasks, dtype=RAY_TASK_DTYPE)
    h = np.zeros(t.shape[0], RAY_HIT_DTYPE)
    check(load().mb3d_hip_march_rays(C.byref(p), _p(t), t.shape[0], int(depth_sqrt),
                                     C.byref(opts) if opts is not None else None, _p(h), 0), "mb3d_hip_march_rays")
    return h


def ray_opts(p: Params, lights=(), positional=(), soft_radius=0.0, hs_mul=1.0, ao_max_len=None) -> "RayOpts":
    """RayOpts: hard shadows of global lights (world directions towards
    them) and positional ones ((scene position, lmax_l)), header indices in
    order; one soft shadow of the last with soft_radius > 0; DE AO at the
    hits with ao_max_len."""
    o = RayOpts()
    i = 0
    for d in lights:
        v = np.asarray(d, np.float64)
        v = -p.step_width * v / np.linalg.norm(v)
        o.light_idx[i], o.light_positional[i] = i, 0
        o.light_vec[i][:] = list(v)
        i += 1
    for pos, lmax in positional:
        o.light_idx[i], o.light_positional[i], o.light_lmax[i] = i, 1, lmax
        o.light_vec[i][:] = list(pos)
        i += 1
    o.n_lights = i
    o.soft_radius, o.hs_mul = soft_radius, hs_mul
    if ao_max_len is not None:
        o.ao, o.ao_max_len = 1, ao_max_len
    return o


def reflection_tasks(p: Params, g: dict, off_steps=2.0):
    """First-bounce reflection rays of a marched picture (march()'s dict
    with depth64 and normals): from each surface pixel, `off_steps` step
    widths off the surface along its normal, the view ray mirrored about
    the normal, one step width long."""
    rays_, step = rays(p)
    return reflection_tasks_from(p, rays_[:, :3], rays_[:, 3:] * step[:, None], g, off_steps)


def reflection_tasks_from(p: Params, starts, vfovs, g: dict, off_steps=2.0):
    """reflection_tasks with the pixels' start points and view vectors (one
    march unit long) given."""
    n = p.width * p.height
    d = np.asarray(g["depth64"], np.float64)
    surf = np.isfinite(d) & (d > 0)
    idx = np.nonzero(surf)[0]
    o = np.asarray(starts, np.float64).reshape(n, 3)[idx]
    v = np.asarray(vfovs, np.float64).reshape(n, 3)[idx]
    vg = np.array([list(r) for r in p.vgrads], np.float64)
    vg /= np.linalg.norm(vg, axis=1, keepdims=True)
    nw = np.asarray(g["normals"], np.float64).reshape(n, 3)[idx] @ vg
    nw /= np.maximum(np.linalg.norm(nw, axis=1, keepdims=True), 1e-300)
    pos = o + v * d[idx][:, None] + nw * (off_steps * p.step_width)
    r = v - 2.0 * np.sum(v * nw, axis=1, keepdims=True) * nw
    t = np.zeros(len(idx), RAY_TASK_DTYPE)
    t["origin"] = pos
    t["dir"] = r.astype(np.float32)
    t["zz"] = d[idx].astype(np.float32)
    t["pixel"] = idx.astype(np.uint32)
    return t


def mc_primary_tasks(p: Params, seed=1):
    """The Monte Carlo renderer's primary rays (CalcRay's first march with
    the inside probe, src/mc.rs:1915-1940) of every pixel: from its start
    plane along its view ray (one march unit long), depth 0, zzplus 1, to
    the far end; seeds from a seeded generator."""
    r, step = rays(p)
    n = r.shape[0]
    t = np.zeros(n, MC_TASK_DTYPE)
    t["pos"] = r[:, :3]
    t["dir"] = r[:, 3:].astype(np.float64) * step[:, None]
    t["vfov"] = t["dir"]
    t["kind"] = MC_RAY
    t["zzplus"] = 1.0
    t["len"] = p.zend
    t["flags"] = MC_PROBE
    t["seed"] = np.random.default_rng(seed).integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    return t


def mc_shadow_tasks(p: Params, res, vfov, light, length=300.0, trans=False):
    """CalcHSMC's shadow rays (src/mc.rs:1513-1566) from the surfaces of a
    batch's results towards a global light (a world direction towards it):
    one step width long, `length` steps at most; the view depth per step
    zz2mul from each result's view vector (`vfov`, one per result)."""
    res = np.asarray(res, MC_RES_DTYPE)
    surf = np.flatnonzero(res["kind"] == 1)
    v = np.asarray(light, np.float64)
    v = v * (p.step_width / np.linalg.norm(v))
    vf = np.asarray(vfov, np.float64)[surf]
    t = np.zeros(len(surf), MC_TASK_DTYPE)
    t["pos"] = res["pos"][surf]
    t["dir"] = v
    t["vfov"] = vf
    t["kind"] = MC_SHADOW
    t["zz"] = t["zz2"] = res["zz"][surf]
    t["zzplus"] = vf @ v / np.sqrt(np.sum(vf * vf, axis=1) * (v @ v))
    t["len"] = length
    t["seed"] = res["seed"][surf]
    t["flags"] = MC_TRANS if trans else 0
    return t, surf


def mc_batch(p: Params, tasks, secant=False):
    """mb3d_hip_mc_batch on host records (an MC_TASK_DTYPE array): the
    results as an MC_RES_DTYPE array."""
    t = np.ascontiguousarray(tasks, dtype=MC_TASK_DTYPE)
    r = np.zeros(t.shape[0], MC_RES_DTYPE)
    check(load().mb3d_hip_mc_batch(C.byref(p), _p(t), t.shape[0], int(secant), _p(r), 0), "mb3d_hip_mc_batch")
    return r


REC_BYTES = 18


def gbuf_pack(gbuf):
    """mb3d_hip_gbuf_pack on host records: (n, 5) u32 -> n * 18 bytes."""
    g = np.ascontiguousarray(gbuf, dtype=np.uint32).reshape(-1, 5)
    out = np.zeros(g.shape[0] * REC_BYTES, np.uint8)
    check(load().mb3d_hip_gbuf_pack(_p(g), g.shape[0], _p(out), 0), "mb3d_hip_gbuf_pack")
    return out


def gbuf_unpack(records):
    """mb3d_hip_gbuf_unpack on host bytes: n * 18 bytes -> (n, 5) u32."""
    r = np.ascontiguousarray(records, dtype=np.uint8).reshape(-1)
    n = r.size // REC_BYTES
    out = np.zeros((n, 5), np.uint32)
    check(load().mb3d_hip_gbuf_unpack(_p(r), n, _p(out), 0), "mb3d_hip_gbuf_unpack")
    return out


# ---------------------------------------------------------------------------
# painting


def lighting_default() -> Lighting:
    """Lighting::default (src/lighting.rs:290-392)."""
    lt = Lighting()
    load().mb3d_hip_lighting_default(C.byref(lt))
    return lt


def paint_setup(s: Scene, p: Params, lt: Lighting = None):
    """(PaintCamera, LightVals) of the scene: paint_camera + light_vals
    (src/render.rs:2340-2380, 2285-2319)."""
    lt = lighting_default() if lt is None else lt
    cam, lv = PaintCamera(), LightVals()
    check(load().mb3d_hip_paint_setup(C.byref(s), C.byref(p), C.byref(lt), C.byref(cam), C.byref(lv)),
          "paint_setup")
    return cam, lv


PAINT_DTYPES = {PAINT_U8: np.uint8, PAINT_U16: np.uint16, PAINT_F32: np.float32}


def paint(cam: PaintCamera, lv: LightVals, gbuf, x0=0, y0=0, w=None, h=None, fmt=PAINT_U8):
    """mb3d_hip_paint on a host G-buffer (n x 5 u32, row-major): the
    picture as an (h, w, 3) array of the format's type."""
    w = cam.width if w is None else w
    h = cam.height if h is None else h
    g = np.ascontiguousarray(gbuf, dtype=np.uint32).reshape(-1, 5)
    if g.shape[0] != w * h:
        raise ValueError(f"the G-buffer holds {g.shape[0]} records, the rectangle {w * h}")
    out = np.zeros((h, w, 3), PAINT_DTYPES[fmt])
    check(load().mb3d_hip_paint(C.byref(cam), C.byref(lv), _p(g), x0, y0, w, h, fmt, _p(out), 0), "mb3d_hip_paint")
    return out


def counters() -> dict:
    c = Counters()
    load().mb3d_hip_counters_get(C.byref(c))
    return {k: getattr(c, k) for k, _ in Counters._fields_}


def chain_surface_point(seed, de, reach=0.1, samples=501):
    """A point of the hybrid chain's surface, where orbits stop escaping
    within the 45 iterations, and the seeded direction `u` it is reached
    from.  The chain's orbits stay bounded only in a core around the origin
    (radius ~0.006-0.008: there the box steps multiply by 8 and the p8 step
    shrinks); its DE shell, 0.07 units further out, holds no bounded orbit.
    A coarse march along the seeded ray from `reach * u` towards the origin
    samples the iteration counts (`de(params, points) -> (de, its)`,
    its = max_it deciding inside); the first inside sample brackets the
    surface, which is bisected to f64 resolution.  Returns (point on the
    escaping side, u)."""
    p = params_from_scene(scene_hybrid(64, 48, seed=seed))
    rng = np.random.default_rng(seed + 3000)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    start = reach * u
    ts = np.linspace(0.0, reach, samples)
    _, its = de(p, start[None, :] - ts[:, None] * u[None, :])
    inside = np.flatnonzero(np.asarray(its) >= p.max_it)
    if inside.size == 0 or inside[0] == 0:
        raise RuntimeError("the coarse march found no surface")
    lo, hi = ts[inside[0] - 1], ts[inside[0]]
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        _, it = de(p, (start - mid * u)[None, :])
        if it[0] >= p.max_it:
            hi = mid
        else:
            lo = mid
    return start - lo * u, u


def scene_deepzoom(width, height, seed, de, distance=1e-9, fov_deg=0.5) -> Scene:
    """The deep zoom on the hybrid chain (Mandelbulb p8 x1, Amazing Box
    scale 2 x2, 45 iterations, numeric DE): the camera's start plane sits
    `distance` scene units outside the surface point of
    `chain_surface_point(seed, de)`, looking back along the seeded ray at
    it, FOV `fov_deg`, the step width matched to the angular pixel
    footprint there (1e-9 at 1080 rows: 8.1e-15 units, ~9,000 f64 ulps of
    the coordinates and 1/58,000 of an FP32 ulp)."""
    surface, u = chain_surface_point(seed, de)
    cam = surface + distance * u
    s = scene_hybrid(width, height, seed=seed)
    s.fov_y = fov_deg
    sw = distance * math.radians(fov_deg) / height
    s.zoom = 2.1345 / (width * sw)
    s.mid[:] = [float(c) for c in cam]
    s.z_start = float(cam[2])  # CalcParams::new offsets the start plane by z_start - mid.z
    s.z_end = s.z_start + 4.0 * distance
    set_camera(s, look_matrix(-u))
    return s


def scene_bulb_deepzoom(width, height, seed, marcher, de, distance=1e-9, fov_deg=0.5,
                        de_mode=DE_ANALYTIC_P8) -> Scene:
    """A deep zoom on the power-8 Mandelbulb: a coarse 64 x 48 march of the
    seeded bulb camera picks a surface pixel (seeded); along its ray the
    first point where the orbit stops escaping is bracketed and bisected to
    f64 resolution (`de(params, points) -> (de, its)`, its = max_it deciding
    inside); the camera's start plane sits `distance` in front of it, FOV
    `fov_deg`, the step width the angular pixel footprint there."""
    coarse = scene_bulb(64, 48, seed=seed, de_mode=de_mode)
    p = params_from_scene(coarse)
    depth = np.asarray(marcher(p), np.float64)
    hits = np.flatnonzero(np.isfinite(depth) & (depth > 0))
    rng = np.random.default_rng(seed + 2000)
    for _ in range(64):
        i = int(hits[rng.integers(hits.size)])
        x, y = i % 64, i // 64
        vg = np.array([[p.vgrads[r][k] for k in range(3)] for r in range(3)])
        c0 = np.array(list(p.ystart)) + vg[0] * x + vg[1] * y
        cafy = (y / p.height - 0.5) * p.fov_y
        cafx = float(np.float32((np.float32(p.fovx_off) - np.float32(x + 1)) * np.float32(p.fovx_mul)))
        w = np.array([-math.sin(cafx), math.sin(cafy), math.cos(cafy) * math.cos(cafx)])
        w /= np.linalg.norm(w)
        v = w @ vg
        u = v / np.linalg.norm(v)
        t0 = depth[i] * np.linalg.norm(v)
        ts = t0 + np.linspace(-0.05, 0.3, 351)
        _, its = de(p, c0[None, :] + ts[:, None] * u[None, :])
        inside = np.asarray(its) >= p.max_it
        k = np.flatnonzero(~inside[:-1] & inside[1:])
        if k.size:
            break
    else:
        raise RuntimeError("no coarse ray enters the set")
    lo, hi = ts[k[0]], ts[k[0] + 1]
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        _, it = de(p, (c0 + mid * u)[None, :])
        if it[0] >= p.max_it:
            hi = mid
        else:
            lo = mid
    cam = c0 + lo * u - distance * u
    s = scene_bulb(width, height, seed=seed, de_mode=de_mode)
    s.fov_y = fov_deg
    sw = distance * math.radians(fov_deg) / height
    s.zoom = 2.1345 / (width * sw)
    s.mid[:] = [float(c) for c in cam]
    s.z_start = float(cam[2])  # CalcParams::new offsets the start plane by z_start - mid.z
    s.z_end = s.z_start + 4.0 * distance
    set_camera(s, look_matrix(u))
    return s


def de_at(p: Params, points):
    """mb3d_hip_de_at: (DE in world units, iteration counts) at points (n x 3)."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    de = np.zeros(len(pts))
    its = np.zeros(len(pts), np.int32)
    check(load().mb3d_hip_de_at(C.byref(p), _p(pts), len(pts), _p(de), _p(its)), "mb3d_hip_de_at")
    return de, its


def catmull_rom_weights(t: float):
    """catmull_rom_weights (src/anim.rs:692-700)."""
    t2, t3 = t * t, t * t * t
    return (0.5 * (-t3 + 2.0 * t2 - t), 0.5 * (3.0 * t3 - 5.0 * t2 + 2.0),
            0.5 * (-3.0 * t3 + 4.0 * t2 + t), 0.5 * (t3 - t2))


SIZES = ((320, 240), (640, 480), (1280, 720), (1920, 1080))


def animation(n_frames=64, seed=1, keys=6, de_mode=DE_ANALYTIC_P8, sizes=SIZES):
    """The batched animation: frames of sizes drawn uniformly from `sizes`
    along a seeded Catmull-Rom camera path around the power-8 Mandelbulb
    (camera distance 1.8-3.0, looking at the origin).  Returns scenes."""
    rng = np.random.default_rng(seed)
    pos = []
    for _ in range(keys):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos.append(d * rng.uniform(1.8, 3.0))
    scenes = []
    for f in range(n_frames):
        u = f / max(n_frames - 1, 1) * (keys - 3)
        k = min(int(u), keys - 4)
        w = catmull_rom_weights(u - k)
        cam = sum(w[i] * pos[k + i] for i in range(4))
        wd, ht = sizes[int(rng.integers(len(sizes)))]
        s = scene_bulb(wd, ht, seed=seed, de_mode=de_mode)
        dist = float(np.linalg.norm(cam))
        s.mid[:] = [0.0, 0.0, 0.0]
        # the start plane through the camera point, facing the origin
        set_camera(s, look_matrix(-cam))
        s.mid[:] = list(cam)
        s.z_start = float(cam[2])
        s.z_end = s.z_start + dist + 30.0
        s.zoom = 0.8 * 2.5 / dist
        scenes.append(s)
    return scenes


def params_array(scenes):
    arr = (Params * len(scenes))()
    for i, s in enumerate(scenes):
        arr[i] = params_from_scene(s)
    return arr


class Plan:
    """A batch of frames captured into a HIP graph (mb3d_hip_plan_*)."""

    def __init__(self, frames, out: Out, opts: Options = None):
        self.lib = load()
        self.h = C.c_void_p()
        self.frames = frames
        self.out = out
        check(self.lib.mb3d_hip_plan_create(frames, len(frames), C.byref(out),
                                            C.byref(opts) if opts is not None else None, C.byref(self.h)),
              "mb3d_hip_plan_create")

    def launch(self, repeats=1):
        check(self.lib.mb3d_hip_plan_launch(self.h, repeats), "mb3d_hip_plan_launch")
        return counters()

    def close(self):
        if self.h:
            self.lib.mb3d_hip_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(frames, out: Out, opts: Options = None):
    o = opts if opts is not None else options()
    ms = C.c_double()
    check(load().mb3d_hip_autotune(frames, len(frames), C.byref(out), C.byref(o), C.byref(ms)), "mb3d_hip_autotune")
    return o, ms.value


MLP_PARAMS = 13634


def _f32(a, cols):
    a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, cols))
    return a, a.ctypes.data_as(C.c_void_p)


class Mlp:
    """The warm-start network (mb3d_hip_mlp_*): rays are n x 6 (origin,
    unit direction), targets n x 2 (hit distance / 4, hit 0/1)."""

    def __init__(self, seed=0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.mb3d_hip_mlp_create(seed, C.byref(self.h)), "mb3d_hip_mlp_create")

    def params(self, new=None):
        if new is not None:
            a = np.ascontiguousarray(np.asarray(new, np.float32))
            assert a.size == MLP_PARAMS
            check(self.lib.mb3d_hip_mlp_params(self.h, None, a.ctypes.data_as(C.c_void_p), MLP_PARAMS), "mlp_params")
        got = np.empty(MLP_PARAMS, np.float32)
        check(self.lib.mb3d_hip_mlp_params(self.h, got.ctypes.data_as(C.c_void_p), None, MLP_PARAMS), "mlp_params")
        return got

    def grad(self, rays, targets):
        r, rp = _f32(rays, 6)
        t, tp = _f32(targets, 2)
        assert len(r) == len(t)
        loss = np.zeros(1, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        check(self.lib.mb3d_hip_mlp_grad(self.h, rp, tp, len(r), loss.ctypes.data_as(C.c_void_p),
                                         g.ctypes.data_as(C.c_void_p)), "mlp_grad")
        return float(loss[0]), g

    def train(self, rays, targets, steps, batch=4096, lr=2e-3):
        r, rp = _f32(rays, 6)
        t, tp = _f32(targets, 2)
        assert len(r) == len(t)
        curve = np.zeros(steps, np.float32)
        check(self.lib.mb3d_hip_mlp_train(self.h, rp, tp, len(r), steps, batch, lr,
                                          curve.ctypes.data_as(C.c_void_p)), "mlp_train")
        return curve

    def infer(self, rays, mfma=True):
        r, rp = _f32(rays, 6)
        out = np.empty((len(r), 2), np.float32)
        check(self.lib.mb3d_hip_mlp_infer(self.h, rp, len(r), int(bool(mfma)), out.ctypes.data_as(C.c_void_p)),
              "mlp_infer")
        return out

    def close(self):
        if self.h:
            self.lib.mb3d_hip_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rays(p: Params, row0=0, rows=None):
    """The pixels' rays as the network takes them (mb3d_hip_rays): n x 6
    float32 (start point, unit direction) and the march unit's length."""
    rows = p.height - row0 if rows is None else rows
    n = rows * p.width
    r = np.empty((n, 6), np.float32)
    sl = np.empty(n, np.float32)
    check(load().mb3d_hip_rays(C.byref(p), row0, rows, _p(r), _p(sl)), "mb3d_hip_rays")
    return r, sl


def network_rays(r):
    """The network's view of rays (n x 6: start point, unit direction): the
    point of closest approach to the scene's origin and the direction, and
    the distance from the start point to that point.  A camera-independent
    form: two cameras that share a ray show it to the network alike."""
    o = np.asarray(r[:, :3], np.float64)
    d = np.asarray(r[:, 3:6], np.float64)
    s = -(o * d).sum(1)
    return np.concatenate([o + s[:, None] * d, d], 1).astype(np.float32), s


def training_pairs(p: Params, opts: Options = None):
    """(ray, exact hit distance) pairs of every pixel of `p` from the plain
    marcher: rays n x 6 in the network's form (network_rays), targets n x 2
    (the hit's distance past the point of closest approach, in scene units
    / 4; hit)."""
    g = march(p, opts=opts, outputs=("depth64",))
    r, sl = rays(p)
    d = g["depth64"]
    hit = np.isfinite(d) & (d > 0)
    t = np.where(hit, d * sl.astype(np.float64), 0.0)
    nr, s = network_rays(r)
    return nr, np.stack([np.where(hit, (t - s) / 4.0, 0.0), hit.astype(np.float64)], 1).astype(np.float32)


def train_warm_start(params_list, steps=3000, batch=8192, lr=3e-3, seed=0, net=None):
    """A network trained on the plain marcher's pairs of every frame of
    `params_list`; returns (net, loss curve)."""
    rs, ts = zip(*(training_pairs(p) for p in params_list))
    rays_all = np.concatenate(rs)
    tg_all = np.concatenate(ts)
    perm = np.random.default_rng(seed).permutation(len(rays_all))
    net = net if net is not None else Mlp(seed)
    curve = net.train(rays_all[perm], tg_all[perm], steps=steps, batch=batch, lr=lr)
    return net, curve
