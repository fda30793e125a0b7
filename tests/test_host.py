"""Host-side checks of the C-ABI library without a GPU: it loads, exports
every function include/mb3d_hip.h declares, derives the march parameters
exactly as the oracle's restatement of CalcParams::new does, and rejects bad
calls with the documented status and a last-error text."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(mb3d_hip_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = H.load()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(H.SIGNATURES), "the ctypes mirror lists the header's functions"


def test_abi_version_and_struct_sizes():
    assert H.load().mb3d_hip_abi_version() == H.ABI_VERSION
    assert C.sizeof(H.Params) % 8 == 0 and C.sizeof(H.Scene) % 8 == 0


def test_ctypes_layouts_match_the_header(tmp_path):
    """Every struct of the mirror against the C compiler's layout of
    include/mb3d_hip.h: size and the offset of every field."""
    structs = {"mb3d_hip_formula": H.Formula, "mb3d_hip_scene": H.Scene, "mb3d_hip_params": H.Params,
               "mb3d_hip_post_job": H.PostJob, "mb3d_hip_out": H.Out, "mb3d_hip_options": H.Options,
               "mb3d_hip_counters": H.Counters, "mb3d_hip_light": H.Light,
               "mb3d_hip_palette_color": H.PaletteColor, "mb3d_hip_lighting": H.Lighting,
               "mb3d_hip_light_val": H.LightVal, "mb3d_hip_light_vals": H.LightVals,
               "mb3d_hip_paint_camera": H.PaintCamera, "mb3d_hip_ray_opts": H.RayOpts}
    # the task and result records as numpy record types
    dtypes = {"mb3d_hip_ray_task": H.RAY_TASK_DTYPE, "mb3d_hip_ray_hit": H.RAY_HIT_DTYPE,
              "mb3d_hip_mc_task": H.MC_TASK_DTYPE, "mb3d_hip_mc_res": H.MC_RES_DTYPE}
    lines, want = [], []
    for cname, cls in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(C.sizeof(cls))
        for fname, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            want.append(getattr(cls, fname).offset)
    for cname, dt in dtypes.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(dt.itemsize)
        for fname in dt.names:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            want.append(dt.fields[fname][1])
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "mb3d_hip.h"\nint main(void) {\n'
                   + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want


def test_scene_default_matches_python_mirror():
    s = H.Scene()
    H.load().mb3d_hip_scene_default(C.byref(s))
    t = H.scene_default()
    assert bytes(s) == bytes(t)


SCENES = {
    "bulb-analytic-640": lambda: H.scene_bulb(640, 480, seed=1),
    "bulb-numeric-1080": lambda: H.scene_bulb(1920, 1080, seed=2, de_mode=H.DE_REFERENCE),
    "box-1080": lambda: H.scene_box(1920, 1080, seed=1),
    "hybrid-1080": lambda: H.scene_hybrid(1920, 1080, seed=3),
}


def _rotated_quaternion():
    s = H.scene_preset("Quaternion", 320, 240)
    s.rot_4d[:] = [0.3, -0.7, 1.1]
    return s


def _mixed_hybrid():
    s = H.scene_preset("Real Power", 320, 240)
    H.set_formulas(s, [H.formula("Quaternion", 2), H.formula("Tricorn"), H.formula("Folding Int Pow", power=4.6),
                       H.formula("Aexion C", mode=21.4, rot_c=0.6)])
    s.repeat_from = 1
    return s


for _name in H.FORMULA_NAMES:
    SCENES["preset-" + _name] = (lambda n: lambda: H.scene_preset(n, 320, 240))(_name)
SCENES["quaternion-rot4d"] = _rotated_quaternion
SCENES["mixed-hybrid"] = _mixed_hybrid


@pytest.mark.parametrize("name", sorted(SCENES))
def test_params_match_the_oracle_bit_for_bit(name):
    sc = SCENES[name]()
    a = H.params_from_scene(sc)
    b = O.params(sc)
    assert bytes(a) == bytes(b)


def test_formula_defaults_follow_default_for():
    """Formula::default_for (src/formulas.rs:266-306), aliases included."""
    want = {
        "Real Power": (H.FORMULA_REALPOWER, [8.0, -1.0]),
        "quat": (H.FORMULA_QUATERNION, [-1.0, 0.0]),
        "Tricorn": (H.FORMULA_TRICORN, [-2.0, 1.0]),
        "bulbox": (H.FORMULA_BULBOX, [2.0, 0.5, 1.0, 1.0, 2.0, 2.0]),
        "folding_int-pow": (H.FORMULA_FOLDINGINTPOW, [2.0, -1.0, 2.0]),
        " Aexion ": (H.FORMULA_AEXIONC, [8.0, 1.0, 1.0, 1.0, 8.0, 1.0, 0.0, 0.0]),
    }
    for name, (kind, opts) in want.items():
        f = H.formula(name)
        assert f.kind == kind and f.iterations == 1
        assert list(f.options)[: len(opts)] == opts and not any(list(f.options)[len(opts):])
    f = H.formula("Mandelbulb")
    assert (f.kind, f.power, f.z_mul) == (H.FORMULA_INTPOW, 8, -1.0)
    f = H.formula("mandelbox")
    assert (f.kind, f.scale, f.min_r, f.fold) == (H.FORMULA_AMAZINGBOX, 2.0, 0.5, 1.0)
    g = H.Formula()
    assert H.load().mb3d_hip_formula_default(b"Menger3", C.byref(g)) == H.UNSUPPORTED


def test_params_of_the_new_formulas():
    """DE option, 4D mode, DE scale and bailout per formula (src/formulas.rs:
    492-528, src/calc.rs:475); the integer options as set_option keeps them."""
    p = H.params_from_scene(H.scene_preset("Quaternion"))
    assert p.de_option == 4 and p.four_d and p.do_julia and p.d_rstop == 256.0
    assert [list(r) for r in p.smatrix4] == [[1.0 if i == k else 0.0 for k in range(4)] for i in range(4)]
    p = H.params_from_scene(H.scene_preset("Bulbox"))
    assert p.de_option == 0 and not p.four_d and p.d_rstop == 1024.0 ** 2
    de_offset = np.float32(min(np.float32(p.de_stop) * np.float32(0.1), np.float32(0.004)))
    assert p.d_de_scale == np.float32(np.float32(0.2) * de_offset)
    p = H.params_from_scene(_mixed_hybrid())
    assert p.de_option == 0 and not p.four_d  # 4 combined with 0 is the 3D numeric DE
    assert p.slot_power[2] == 5 and p.slot_opt[2][0] == 5.0
    assert list(p.slot_opt[3]) == [8.0, 1.0, 1.0, 1.0, 8.0, 1.0, 0.0, 21.0]


def test_params_of_the_hybrid_chain():
    """p8 x1 + box x2 combine to the numeric DE (src/calc.rs:318-372) with
    dDEscale = (1 * 1.0 + 2 * 0.2) / 3 times the DE offset."""
    p = H.params_from_scene(H.scene_hybrid(640, 480))
    assert p.de_option == 0 and not p.is_custom_de
    de_offset = np.float32(min(np.float32(p.de_stop) * np.float32(0.1), np.float32(0.004)))
    assert p.d_de_scale == np.float32(np.float32(1.4 / 3.0) * de_offset)
    assert p.d_rstop == 1024.0 ** 2


def test_params_of_the_box():
    p = H.params_from_scene(H.scene_box(1920, 1080))
    assert p.de_option == 11 and p.is_custom_de and p.slot_ade[0]
    assert p.d_de_scale == np.float32(np.float64(np.float32(0.65)) / p.step_width)


def test_pixel_seed_matches_oracle():
    lib = H.load()
    for y, x in [(0, 0), (7, 123), (1079, 1919), (1499, 2999)]:
        assert lib.mb3d_hip_pixel_seed(y, x) == O.pixel_seed(y, x)


def test_invalid_scenes_are_rejected():
    lib = H.load()
    s = H.scene_default()
    s.n_formulas = 0
    p = H.Params()
    assert lib.mb3d_hip_params_from_scene(C.byref(s), C.byref(p)) == -1
    assert "formula" in H.last_error()
    s = H.scene_default()
    s.formulas[0].iterations = 0
    assert lib.mb3d_hip_params_from_scene(C.byref(s), C.byref(p)) == -1
    s = H.scene_default()
    s.abi_version = 99
    assert lib.mb3d_hip_params_from_scene(C.byref(s), C.byref(p)) == -1
    assert lib.mb3d_hip_params_from_scene(None, C.byref(p)) == -1


def test_unsupported_scenes_name_the_reason():
    lib = H.load()
    s = H.scene_bulb(64, 48)  # the analytic power-8 DE in Julia mode
    s.julia = 1
    p = H.params_from_scene(s)
    buf = C.create_string_buffer(200)
    assert lib.mb3d_hip_supported(C.byref(p), buf, 200) == H.UNSUPPORTED
    assert b"Julia" in buf.value
    s = H.scene_default()
    s.formulas[0].power = 5  # other powers march in FP64
    p = H.params_from_scene(s)
    assert lib.mb3d_hip_supported(C.byref(p), buf, 200) == H.OK
    s = H.scene_default()
    s.smooth_normals = 3  # smooth normals, the roughness: the general normals pass
    p = H.params_from_scene(s)
    assert lib.mb3d_hip_supported(C.byref(p), buf, 200) == H.OK
    # a colour option of a DE combination's ruling part: not on this backend
    s = H.scene_preset("Integer Power", 64, 48)
    H.set_formulas(s, [H.formula("Integer Power"), H.formula("Amazing Box", scale=-1.5)])
    s.decomb, s.decomb_end1, s.decomb_start2, s.decomb_end2, s.decomb_repeat2 = 1, 0, 1, 1, 1
    s.decomb_iterations2 = 20
    s.color_option = 2
    p = H.params_from_scene(s)
    assert lib.mb3d_hip_supported(C.byref(p), buf, 200) == H.UNSUPPORTED
    assert b"DE combination" in buf.value
    p = H.params_from_scene(H.scene_bulb(64, 48))
    assert lib.mb3d_hip_supported(C.byref(p), buf, 200) == H.OK


def test_march_before_init_fails_without_touching_a_device():
    lib = H.load()
    p = H.params_from_scene(H.scene_bulb(64, 48))
    o = H.Out()
    st = lib.mb3d_hip_march(C.byref(p), None, 0, 48, C.byref(o), None, H.PROGRESS_FN(0), H.CANCEL_FN(0), None)
    assert st == -3 and "init" in H.last_error()


def _ipol_scene(f0, f1, w, four=False):
    s = H.scene_preset("Quaternion" if four else "Integer Power", 320, 240)
    H.set_formulas(s, [f0, f1])
    s.interpolation = 1
    s.interpolation_w[:] = w
    return s


@pytest.mark.parametrize("case", ["p8-box", "box-box-analytic", "quaternions-4d", "negative-weight"])
def test_interpolation_params_match_the_oracle(case):
    sc = {
        "p8-box": lambda: _ipol_scene(H.formula("Integer Power"), H.formula("Amazing Box"), [0.7, 0.3]),
        "box-box-analytic": lambda: _ipol_scene(H.formula("Amazing Box"), H.formula("Amazing Box", scale=-1.5),
                                                [2.0, 1.0]),
        "quaternions-4d": lambda: _ipol_scene(H.formula("Quaternion"), H.formula("Quaternion", yw_mul=0.5),
                                              [1.0, 3.0], four=True),
        "negative-weight": lambda: _ipol_scene(H.formula("Integer Power"), H.formula("Amazing Box"), [-0.5, 1.0]),
    }[case]()
    a = H.params_from_scene(sc)
    b = O.params(sc)
    assert a.ipol and bytes(a) == bytes(b)


def _decomb_scene(kind, place=False, f2=None, its2=20):
    s = H.scene_preset("Integer Power", 320, 240)
    s.z_start, s.zoom = -4.0, 0.45
    H.set_formulas(s, [H.formula("Integer Power"), H.formula("Real Power", 2), f2 or H.formula("Amazing Box", scale=-1.5),
                       H.formula("Tricorn", -1)])
    s.decomb, s.decomb_end1, s.decomb_start2, s.decomb_end2, s.decomb_repeat2 = kind, 1, 2, 3, 3
    s.decomb_iterations2, s.decomb_smooth = its2, 0.3
    if place:
        s.part2_placed = 1
        s.part2_offset[:], s.part2_rotation[:], s.part2_scale = (1.6, 0.3, 0.0), (10.0, 20.0, 30.0), 0.35
    return s


@pytest.mark.parametrize("case", ["min", "max-placed", "inverted", "smooth-linear", "smooth", "4d-second",
                                  "uncounted-second"])
def test_decomb_params_match_the_oracle(case):
    sc = {
        "min": lambda: _decomb_scene(1),
        "max-placed": lambda: _decomb_scene(2, place=True),
        "inverted": lambda: _decomb_scene(3),
        "smooth-linear": lambda: _decomb_scene(4),
        "smooth": lambda: _decomb_scene(5, place=True, its2=40),
        "4d-second": lambda: _decomb_scene(1, f2=H.formula("Quaternion", 3)),
        "uncounted-second": lambda: _decomb_scene(1, f2=H.formula("Bulbox", -2)),
    }[case]()
    a = H.params_from_scene(sc)
    b = O.params(sc)
    assert a.decomb and bytes(a) == bytes(b)
    buf = C.create_string_buffer(200)
    assert H.load().mb3d_hip_supported(C.byref(a), buf, 200) == H.OK, buf.value
