"""The oracle's DE combinations (calc_de_fractal, src/calc.rs:1352-1456)
pinned by what they are built from: each part's DE is that of a picture of
the part alone (same bailout, the second part's iteration limit), so

  kind 1 (min)                      DE = min(DE1, DE2), the ruling part's iterations
  kind 2 (max)                      DE = max(DE1, DE2)
  kind 3 (max, part 2 inverted)     DE = max(DE1, 4 stop - 3 DE2)
  kinds 4, 5 (smooth minima)        <= min(DE1, DE2), = min(DE1, DE2) where both are
                                    beyond the smoothing distance
  placement                         part 2 at m (p - offset), distances times the scale
  (CPU only)."""
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

RSTOP = 16.0  # one bailout for the combination and the parts alone


def combo(kind, place=None, f2=None, its2=20, smooth=0.3):
    s = H.scene_preset("Integer Power", 160, 120)
    s.z_start, s.zoom = -4.0, 0.45
    s.rstop = RSTOP
    H.set_formulas(s, [H.formula("Integer Power"), f2 or H.formula("Amazing Box", scale=-1.5)])
    s.decomb, s.decomb_end1, s.decomb_start2, s.decomb_end2, s.decomb_repeat2 = kind, 0, 1, 1, 1
    s.decomb_iterations2, s.decomb_smooth = its2, smooth
    if place:
        s.part2_placed = 1
        s.part2_offset[:], s.part2_rotation[:], s.part2_scale = place
    return s


def alone(f, its):
    s = H.scene_preset("Integer Power", 160, 120)
    s.z_start, s.zoom = -4.0, 0.45
    s.rstop = RSTOP
    s.iterations = its
    H.set_formulas(s, [f])
    return O.params(s)


POINTS = np.random.default_rng(7).uniform(-2.2, 2.2, (300, 3))


def parts(f2=None, its2=20):
    return alone(H.formula("Integer Power"), 12), alone(f2 or H.formula("Amazing Box", scale=-1.5), its2)


def test_params_of_a_combination():
    p = O.params(combo(1))
    assert (p.decomb, p.end_to, p.start2, p.end2, p.repeat2, p.max_it2) == (1, 0, 1, 1, 1, 20)
    assert p.de_option == 0 and not p.is_custom_de and p.de_option2 == 11 and p.is_custom_de2
    assert list(p.slot_ade[:2]) == [0, 1]
    a, b = parts()
    assert p.d_de_scale == a.d_de_scale and p.d_de_scale2 == b.d_de_scale
    # the colour scale counts the larger iteration limit (iMaxitF2)
    assert p.mcts_m == np.float32(32767.0) / np.float32(21.0)
    # no second part with iterations: no combination
    s = combo(1)
    s.formulas[1].iterations = 0
    assert O.params(s).decomb == 0
    with pytest.raises(RuntimeError, match="dIFS"):
        O.params(combo(6))


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_min_max_and_inverted_max(kind):
    p = O.params(combo(kind))
    a, b = parts()
    ms = float(np.float32(p.de_stop))
    for q in POINTS:
        d, its = O.de_at(p, q)
        d1, i1 = O.de_at(a, q)
        d2, i2 = O.de_at(b, q)
        d1, d2 = d1 / a.step_width, d2 / b.step_width
        d /= p.step_width
        if kind == 1:
            want, want_its = (d2, i2) if d2 < d1 else (d1, i1)
        elif kind == 2:
            want, want_its = (d2, i2) if d2 > d1 else (d1, i1)
        else:
            inv = ms * 4.0 - d2 * 3.0
            want, want_its = (inv, i2 - (inv >= ms)) if inv > d1 else (d1, i1)
        assert d == pytest.approx(want, rel=1e-12, abs=1e-12), q
        assert its == want_its, q


@pytest.mark.parametrize("kind", [4, 5])
def test_smooth_minima(kind):
    p = O.params(combo(kind, smooth=0.02))
    a, b = parts()
    sm = p.decomb_smooth
    far = 0
    for q in POINTS:
        d = O.de_at(p, q)[0] / p.step_width
        d1 = O.de_at(a, q)[0] / a.step_width
        d2 = O.de_at(b, q)[0] / b.step_width
        assert d <= min(d1, d2) + 1e-9
        if min(d1, d2) >= sm:
            far += 1
            assert d == pytest.approx(min(d1, d2), rel=1e-12)
    assert far > 40


def test_placed_second_part():
    """A translated, rotated and scaled second part: its DE is the scale
    times the part's own DE at m (p - offset), m the rotation divided by
    the scale (build_rot_matrix, src/math.rs:466-475)."""
    off, rot, scale = (1.6, 0.3, 0.0), (10.0, 20.0, 30.0), 0.35
    p = O.params(combo(1, place=(off, rot, scale)))
    assert p.placed and p.place_scale == scale
    a, b = parts()
    xa, ya, za = (math.radians(r) for r in rot)
    sx, cx, sy, cy, sz, cz = math.sin(xa), math.cos(xa), math.sin(ya), math.cos(ya), math.sin(za), math.cos(za)
    r = np.array([[cy * cz, -cy * sz, sy], [sx * sy * cz + cx * sz, cx * cz - sx * sy * sz, -sx * cy],
                  [sx * sz - cx * sy * cz, cx * sy * sz + sx * cz, cx * cy]])
    assert np.allclose(np.array(p.place_m), r / scale, rtol=1e-14)
    assert np.allclose(r @ r.T, np.eye(3))
    for q in POINTS[:120]:
        d = O.de_at(p, q)[0] / p.step_width
        d1 = O.de_at(a, q)[0] / a.step_width
        d2 = O.de_at(b, (r / scale) @ (q - np.array(off)))[0] / b.step_width * scale
        assert d == pytest.approx(min(d1, d2), rel=1e-9, abs=1e-9)


def test_4d_second_part():
    """A Quaternion second part runs the 4D hybrid inside the combination."""
    f2 = H.formula("Quaternion")
    p = O.params(combo(1, f2=f2, its2=12))
    assert p.four_d2 and p.de_option2 == 4 and not p.four_d
    a, b = parts(f2, 12)
    for q in POINTS[:100]:
        d = O.de_at(p, q)[0] / p.step_width
        d1 = O.de_at(a, q)[0] / a.step_width
        d2 = O.de_at(b, q)[0] / b.step_width
        assert d == pytest.approx(min(d1, d2), rel=1e-12)


@pytest.mark.parametrize("kind", [1, 2, 4, 5])
def test_combinations_render(kind):
    place = ((1.6, 0.3, 0.0), (10.0, 20.0, 30.0), 0.35) if kind == 2 else None
    r = O.render(O.params(combo(kind, place=place)), threads=4)
    assert 0.005 < np.isfinite(r["depth"]).mean() < 0.995
