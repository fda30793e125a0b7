"""The built-in formulas beyond power-8 Integer Power and Amazing Box on the
HIP backend (FP64 march, kind K_GEN) against the f64 oracle: every preset
of the reference (Scene::preset, src/scene.rs:811-838), the other integer
powers, a rotated 4D Quaternion, a mixed hybrid, Julia modes, and the post
passes.  The bar is the parity bar of tests/test_gpu_parity.py; the
formulas with libm transcendentals (Real Power, Aexion C) differ from the
oracle by the device's sin, cos, atan2, pow, exp and log (within an ulp or
two), the others follow the oracle's operation order exactly."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu


def run(hip, sc, scale, opts=None, post=None):
    p = hip.params_from_scene(sc)
    g = hip.march(p, post=post, opts=opts)
    r = O.render(O.params(sc), post, threads=0)
    m = parity.compare(g, r, p.width, p.height, p.step_width, scale)
    print(f"{sc.width}x{sc.height}", {k: m[k] for k in ("mask_agree", "depth_rel_p99", "normal_deg_p99",
                                                     "iters_equal_inner", "both")}, flush=True)
    return p, g, r, m


def check(m, iters_eq=0.995):
    assert m["mask_agree"] >= 0.999, m
    assert m["depth_ok_frac"] >= 0.999 and m["depth_rel_p99"] <= 1e-4, m
    assert m["normal_ok_frac"] >= 0.999 and m["normal_deg_p99"] < 0.5, m
    assert m["iters_equal_inner"] >= iters_eq, m


# the scene scale of each preset: the radius the object fits in
SCALE = {"Amazing Box": 4.0, "Bulbox": 4.0, "Folding Int Pow": 4.0}


@pytest.mark.parametrize("name", sorted(H.FORMULA_NAMES))
def test_every_preset_matches_the_oracle(hip, name):
    sc = H.scene_preset(name, 640, 480)
    p, g, r, m = run(hip, sc, SCALE.get(name, 2.0))
    check(m)
    c = g["counters"]
    if name not in ("Integer Power", "Amazing Box"):
        assert c["de_evals_f32"] == 0 and c["its_other_f64"] > 0, c
        # the work the oracle did, within the share of pixels that differ
        oc = r["counters"]["its_other"]
        assert abs(c["its_other_f64"] - oc) <= 0.05 * oc, (c["its_other_f64"], oc)


@pytest.mark.parametrize("power", [2, 3, 4, 5, 6, 7])
def test_integer_powers(hip, power):
    sc = H.scene_preset("Integer Power", 320, 240)
    sc.iterations = 20
    H.set_formulas(sc, [H.formula("Integer Power", power=power)])
    _, g, _, m = run(hip, sc, 2.0)
    check(m)
    assert g["counters"]["its_p8_f64"] > 0 and g["counters"]["de_evals_f32"] == 0


def test_rotated_quaternion(hip):
    sc = H.scene_preset("Quaternion", 480, 360)
    sc.rot_4d[:] = [0.3, -0.5, 0.8]
    _, _, r, m = run(hip, sc, 2.0)
    assert np.isfinite(r["depth"]).mean() > 0.05
    check(m)


def test_quaternion_mandelbrot_slice(hip):
    """The Quaternion outside Julia mode: the 3D slice of the 4D set."""
    sc = H.scene_preset("Quaternion", 480, 360)
    sc.julia = 0
    sc.z_start, sc.zoom = -3.5, 0.6
    _, _, r, m = run(hip, sc, 2.0)
    assert np.isfinite(r["depth"]).mean() > 0.05
    check(m)


def test_mixed_hybrid(hip):
    """Quaternion x2, Tricorn, Folding Int Pow 5, Aexion C mode 21 (without
    turning C), repeating from the second slot: the 3D numeric DE."""
    sc = H.scene_preset("Real Power", 480, 360)
    H.set_formulas(sc, [H.formula("Quaternion", 2), H.formula("Tricorn"), H.formula("Folding Int Pow", power=5),
                        H.formula("Aexion C", mode=21, rot_c=0)])
    sc.repeat_from = 1
    p, _, r, m = run(hip, sc, 2.0)
    assert not p.four_d and np.isfinite(r["depth"]).mean() > 0.02
    check(m)


def test_julia_bulb_and_box(hip):
    """Julia mode with the built-in kernels' formulas: the power-8 bulb
    (numeric DE) and the Amazing Box (analytic DE) both go to the general
    FP64 path."""
    sc = H.scene_preset("Integer Power", 480, 360)
    sc.julia, sc.iterations = 1, 10
    sc.julia_c[:] = [0.35, -0.2, 0.15, 0.0]
    _, g, r, m = run(hip, sc, 2.0)
    assert np.isfinite(r["depth"]).mean() > 0.05
    check(m)
    sc = H.scene_preset("Amazing Box", 480, 360)
    sc.julia = 1
    sc.julia_c[:] = [0.8, -0.6, 0.4, 0.0]
    p, g, r, m = run(hip, sc, 4.0)
    assert p.de_option == 11 and np.isfinite(r["depth"]).mean() > 0.05
    check(m)


def test_fp32_request_marches_in_fp64(hip):
    """The general formulas have no FP32 iterate: an FP32 request marches
    them in FP64 (and says so in the counters)."""
    sc = H.scene_preset("Tricorn", 320, 240)
    p = hip.params_from_scene(sc)
    a = hip.march(p)
    b = hip.march(p, opts=H.options(precision=H.PREC_F32))
    assert b["counters"]["de_evals_f32"] == 0
    assert np.array_equal(a["gbuf"], b["gbuf"])


@pytest.mark.parametrize("name", ["Real Power", "Quaternion", "Bulbox", "Aexion C"])
def test_de_at_matches_the_oracle(hip, name):
    """The f64 DE at random points (mb3d_hip_de_at): iteration counts equal
    at >= 99 % of them, the DE within 1e-9 relative at >= 99 %."""
    sc = H.scene_preset(name, 320, 240)
    p = hip.params_from_scene(sc)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1.6, 1.6, (4000, 3))
    de, its = hip.de_at(p, pts)
    po = O.params(sc)
    ref = [O.de_at(po, q) for q in pts]
    de_o = np.array([a for a, _ in ref])
    its_o = np.array([b for _, b in ref])
    assert (its == its_o).mean() >= 0.99
    rel = np.abs(de - de_o) / np.maximum(np.abs(de_o), 1e-300)
    assert (rel <= 1e-9).mean() >= 0.99, np.percentile(rel, [50, 99, 100])


def test_post_passes_on_a_general_formula(hip):
    """Hard shadows and DE AO on the Quaternion preset (the f64 post
    kernels of the general kind)."""
    sc = H.scene_preset("Quaternion", 320, 240)
    p = hip.params_from_scene(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], deao_quality=1, params=p)
    g = hip.march(p, post=job)
    r = O.render(O.params(sc), job, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0) & np.isfinite(g["depth64"]) & (g["depth64"] > 0)
    assert surf.sum() > 1000
    sg = (g["gbuf"][surf, 1] >> 16) & 0xFC00
    so = (r["gbuf"][surf, 1] >> 16) & 0xFC00
    assert (sg == so).mean() >= 0.99 and (so != 0).any()
    ag = (g["gbuf"][surf, 4] & 0xFFFF).astype(np.float64) / 16383.0
    ao = (r["gbuf"][surf, 4] & 0xFFFF).astype(np.float64) / 16383.0
    assert np.abs(ag - ao).mean() < 0.01


def test_batches_reject_the_general_formulas(hip):
    """The batched plans run the FP32 march: a general formula's frames are
    refused with a reason, not marched otherwise."""
    sc = H.scene_preset("Tricorn", 64, 48)
    p = hip.params_from_scene(sc)
    arr = H.params_array([sc])
    out = H.Out()
    out.on_device = 1
    st = H.load().mb3d_hip_march_batch(arr, 1, H.C.byref(out), None)
    assert st == H.UNSUPPORTED and "FP32" in H.last_error()
    assert p.width == 64


def ipol_scene(f0, f1, w, preset, size=(480, 360)):
    s = H.scene_preset(preset, *size)
    H.set_formulas(s, [f0, f1])
    s.interpolation = 1
    s.interpolation_w[:] = w
    return s


@pytest.mark.parametrize("case", ["real-tricorn", "p8-box", "box-box-analytic", "quaternions-4d"])
def test_interpolation_hybrids(hip, case):
    """The interpolation hybrids (run_ipol, src/iteration.rs:347-401): 3D
    numeric, 3D with the box's analytic DE, 4D numeric."""
    sc, scale = {
        "real-tricorn": (ipol_scene(H.formula("Real Power", power=5.5), H.formula("Tricorn"), [0.5, 0.5],
                                    "Integer Power"), 2.0),
        "p8-box": (ipol_scene(H.formula("Integer Power"), H.formula("Amazing Box"), [0.7, 0.3], "Integer Power"),
                   2.0),
        "box-box-analytic": (ipol_scene(H.formula("Amazing Box"), H.formula("Amazing Box", scale=-1.5), [2.0, 1.0],
                                        "Amazing Box"), 4.0),
        "quaternions-4d": (ipol_scene(H.formula("Quaternion"), H.formula("Quaternion", yw_mul=0.5), [1.0, 3.0],
                                      "Quaternion"), 2.0),
    }[case]
    p, g, r, m = run(hip, sc, scale)
    assert p.ipol and np.isfinite(r["depth"]).mean() > 0.01
    check(m)
