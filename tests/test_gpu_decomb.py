"""DE combinations (calc_de_fractal, src/calc.rs:1352-1456) and uncounted
hybrid slots on the HIP backend (the general FP64 kind) against the f64
oracle, at the parity bar of tests/test_gpu_parity.py."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu

PLACE = ((1.6, 0.3, 0.0), (10.0, 20.0, 30.0), 0.35)


def combo(kind, place=None, f2=None, its2=20, size=(480, 360)):
    s = H.scene_preset("Integer Power", *size)
    s.z_start, s.zoom = -4.0, 0.45
    H.set_formulas(s, [H.formula("Integer Power"), f2 or H.formula("Amazing Box", scale=-1.5)])
    s.decomb, s.decomb_end1, s.decomb_start2, s.decomb_end2, s.decomb_repeat2 = kind, 0, 1, 1, 1
    s.decomb_iterations2, s.decomb_smooth = its2, 0.3
    if place:
        s.part2_placed = 1
        s.part2_offset[:], s.part2_rotation[:], s.part2_scale = place
    return s


def run(hip, sc, scale=2.0, post=None):
    p = hip.params_from_scene(sc)
    g = hip.march(p, post=post)
    r = O.render(O.params(sc), post, threads=0)
    m = parity.compare(g, r, p.width, p.height, p.step_width, scale)
    print({k: m[k] for k in ("mask_agree", "depth_rel_p99", "normal_deg_p99", "iters_equal_inner", "both")},
          flush=True)
    return p, g, r, m


def check(m, iters_eq=0.995):
    assert m["mask_agree"] >= 0.999, m
    assert m["depth_ok_frac"] >= 0.999 and m["depth_rel_p99"] <= 1e-4, m
    assert m["normal_ok_frac"] >= 0.999 and m["normal_deg_p99"] < 0.5, m
    assert m["iters_equal_inner"] >= iters_eq, m


@pytest.mark.parametrize("kind,placed", [(1, False), (1, True), (2, True), (3, True), (4, False), (5, True)])
def test_combinations_match_the_oracle(hip, kind, placed):
    p, g, r, m = run(hip, combo(kind, PLACE if placed else None))
    assert p.decomb == kind and np.isfinite(r["depth"]).mean() > 0.005
    assert g["counters"]["de_evals_f32"] == 0
    check(m)


def test_combination_with_a_4d_second_part(hip):
    p, _, r, m = run(hip, combo(1, PLACE, f2=H.formula("Quaternion"), its2=12))
    assert p.four_d2 and np.isfinite(r["depth"]).mean() > 0.005
    check(m)


def test_combination_gbuffer_colour_and_post(hip):
    """The G-buffer words (colour from the ruling part's trap and smooth
    iterations) and a shadow pass on a smooth combination."""
    sc = combo(5, PLACE, size=(320, 240))
    p = hip.params_from_scene(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], params=p)
    g = hip.march(p, post=job)
    r = O.render(O.params(sc), job, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0) & np.isfinite(g["depth64"]) & (g["depth64"] > 0)
    assert surf.sum() > 1000
    for w in (1, 3):  # shadow bits and si_gradient / otrap
        assert (g["gbuf"][surf, w] == r["gbuf"][surf, w]).mean() >= 0.99, w


def test_uncounted_slots(hip):
    """A slot with negative iterations is iterated but not counted
    (src/iteration.rs:294-300)."""
    s = H.scene_preset("Integer Power", 480, 360)
    H.set_formulas(s, [H.formula("Integer Power"), H.formula("Tricorn", -1)])
    p, g, r, m = run(hip, s)
    assert p.slot_uncounted[1] and np.isfinite(r["depth"]).mean() > 0.02
    check(m)
