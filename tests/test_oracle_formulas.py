"""The oracle's built-in formulas beyond Integer Power 8 and Amazing Box,
pinned by the reference's own known-answer tests and by what the
mathematics fixes (CPU only).

Reference tests re-expressed here:
  src/formulas.rs:944-955  Real Power 8 = Integer Power 8 (1e-9)
  src/formulas.rs:881-915  Aexion C against the Pascal reference (mode 0)
  tests/render.rs:24-31    every preset hits the object (tests/test_oracle.py)
Closed forms and reductions:
  Real Power n = Integer Power n for n = 2..7 (the spherical triplex)
  Aexion C, power 2, C not rotated = the y-pole triplex square (a polynomial)
  Aexion C mode 8 = mode 0 with C's y and z swapped (the new constant)
  Quaternion and Tricorn in the complex plane = the Mandelbrot iteration
  Bulbox below R threshold 2 = Integer Power 2 (bulb scaling 1); above R
  threshold = Amazing Box; between, the length blends linearly
  Folding Int Pow with a fold beyond the point = Integer Power
  build_smatrix4: orthonormal, single-plane rotations by hand
  interpolation hybrids: F with itself = F; (F0, F1, w0, w1) = (F1, F0, w1, w0)
"""
import cmath
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

POINTS = [(0.31, -0.42, 0.27), (-0.7, 0.2, 0.1), (0.05, 0.9, -0.2), (0.6, 0.55, -0.45)]


def norm2(p):
    return float(sum(a * a for a in p[:3]))


@pytest.mark.parametrize("n", range(2, 9))
def test_real_power_equals_integer_power(n):
    """src/formulas.rs:944-955 for n = 8; the other integer powers follow
    from the same spherical triplex formula."""
    for p in POINTS:
        v = [p[0], p[1], p[2], 0.0]
        a, _ = O.iterate(H.formula("Integer Power", power=n, z_mul=-1.0), v, [0.0] * 4)
        b, _ = O.iterate(H.formula("Real Power", power=float(n), z_mul=-1.0), v, [0.0] * 4, rout=norm2(p))
        assert np.abs(a[:3] - b[:3]).max() < 1e-9 * max(1.0, np.abs(a).max()), (n, p)


def test_real_power_reads_the_stored_radius():
    """The radius comes from `rout`, not from the vector (HybridFloatPow)."""
    f = H.formula("Real Power", power=3.5)
    v = [0.3, 0.2, -0.1, 0.0]
    a, _ = O.iterate(f, v, [0.0] * 4, rout=0.25)
    b, _ = O.iterate(f, v, [0.0] * 4, rout=1.0)
    assert np.allclose(a[:3] / b[:3], 0.25 ** 1.75, rtol=1e-12)


def pascal_aexion(x, y, z, cx, cy, cz):
    """The reference's Pascal restatement of Aexion C (src/formulas.rs:881-915):
    power 8, power C 8, rotate C and conditional phi on, mode 0."""
    p = 8.0
    r1 = x * x + y * y + z * z
    th = math.atan2(math.sqrt(x * x + z * z), y) * p
    ph = math.atan2(z, x) * p
    r1 = r1 ** (p * 0.5)
    xn, yn, zn = math.cos(th) * math.cos(ph) * r1 + cx, math.cos(th) * math.sin(ph) * r1 + cy, r1 * math.sin(th) + cz
    r = math.sqrt(cx * cx + cy * cy + cz * cz)
    th2 = math.atan2(math.sqrt(cx * cx + cz * cz), cy) * 8.0
    ph2 = math.atan2(cz, cx) * 8.0
    if xn > 0.0:
        ph2 = -ph2
    return (xn, yn, zn), (math.cos(th2) * math.cos(ph2) * r, math.cos(th2) * math.sin(ph2) * r, r * math.sin(th2))


def test_aexion_c_matches_the_pascal_reference():
    f = H.formula("Aexion C")
    for x, y, z, cx, cy, cz in [(0.3, -0.4, 0.5, 0.2, 0.1, -0.3), (-0.7, 0.2, 0.1, -0.5, 0.4, 0.6),
                                (0.05, 0.9, -0.2, 0.3, -0.8, 0.1)]:
        v, j = O.iterate(f, [x, y, z, 0.0], [cx, cy, cz, 0.0])
        vr, jr = pascal_aexion(x, y, z, cx, cy, cz)
        assert np.abs(v[:3] - vr).max() < 1e-9
        assert np.abs(j[:3] - jr).max() < 1e-9


def test_aexion_c_power_two_is_the_y_pole_triplex_square():
    """rotate C off, power 2: r^2 (cos 2t cos 2p, cos 2t sin 2p, z_mul sin 2t)
    with t from the y axis and p = atan2(z, x), as polynomials."""
    f = H.formula("Aexion C", power=2.0, z_mul=-1.3, rot_c=0.0)
    for x, y, z in POINTS:
        q = x * x + z * z
        want = [(y * y - q) * (x * x - z * z) / q, (y * y - q) * 2 * x * z / q, -1.3 * 2 * y * math.sqrt(q)]
        c = [0.1, -0.2, 0.3]
        v, j = O.iterate(f, [x, y, z, 0.0], c + [0.0])
        assert np.abs(v[:3] - (np.array(want) + c)).max() < 1e-12
        assert list(j[:3]) == c  # the constant is not turned


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_aexion_c_mode_8_swaps_the_constants_y_and_z(mode):
    """With conditional phi off, mode bit 8 only swaps the two components the
    angles are taken from: the turned constant equals that of the same mode
    with C's y and z exchanged."""
    for x, y, z in POINTS:
        c = [0.2, -0.35, 0.6]
        cs = [c[0], c[2], c[1]]
        _, j8 = O.iterate(H.formula("Aexion C", cond_phi=0.0, mode=mode + 8), [x, y, z, 0.0], c + [0.0])
        _, j0 = O.iterate(H.formula("Aexion C", cond_phi=0.0, mode=mode), [x, y, z, 0.0], cs + [0.0])
        assert np.abs(j8[:3] - j0[:3]).max() < 1e-12


def mandel_count(c, max_it, rstop2):
    z = c
    for it in range(1, max_it + 1):
        z = z * z + c
        if abs(z) ** 2 > rstop2:
            return it
    return max_it


@pytest.mark.parametrize("name", ["Quaternion", "Tricorn"])
def test_complex_plane_reduces_to_the_mandelbrot_set(name):
    """Points of the xy plane (z = w = 0) stay there under Quaternion and
    Tricorn, which then square like complex numbers: the iteration count at
    such a point is the Mandelbrot escape count."""
    s = H.scene_default()
    s.iterations = 40
    H.set_formulas(s, [H.formula(name)])
    p = O.params(s)
    rng = np.random.default_rng(3)
    for _ in range(60):
        x, y = rng.uniform(-2.0, 0.6), rng.uniform(-1.2, 1.2)
        _, its = O.de_at(p, [x, y, 0.0])
        assert its == mandel_count(complex(x, y), 40, p.d_rstop), (x, y)


def test_tricorn_by_hand():
    v, _ = O.iterate(H.formula("Tricorn"), [0.5, 0.25, 0.5, 0.0], [0.1, 0.2, 0.3, 0.0])
    assert np.allclose(v[:3], [0.25 - (0.0625 + 0.25) + 0.1, 2 * 0.125 + 0.2, 0.5 * 0.5 * -2.0 + 0.3], atol=1e-15)


def test_quaternion_by_hand():
    f = H.formula("Quaternion", yw_mul=0.5, w_add=0.25)
    x, y, z, w = 0.3, -0.2, 0.4, 0.1
    j = [0.01, 0.02, 0.03, 0.04]
    v, _ = O.iterate(f, [x, y, z, w], j)
    want = [x * x - y * y - z * z - w * w + j[0], 2 * (y * x + z * w) + j[1], 2 * (z * x + 0.5 * y * w) + j[2],
            2 * (w * x + y * z) + 0.25 + j[3]]
    assert np.allclose(v, want, atol=1e-15)


def test_bulbox_regions():
    f = H.formula("Bulbox", r_threshold=1.0, r_threshold2=0.5)
    j = [0.05, -0.1, 0.2, 0.0]
    bulb = H.formula("Integer Power", power=2, z_mul=-1.0)
    box = H.formula("Amazing Box")
    v = [0.3, -0.2, 0.25, 0.0]
    # inside R threshold 2: the power-2 bulb (bulb scaling 1 is exact)
    a, _ = O.iterate(f, v, j, rout=0.2)
    b, _ = O.iterate(bulb, v, j)
    assert np.array_equal(a[:3], b[:3])
    # outside R threshold: the box
    a, _ = O.iterate(f, v, j, rout=1.5)
    b, _ = O.iterate(box, v, j)
    assert np.array_equal(a[:3], b[:3])
    # between: the length is the blend of the two lengths, the direction that
    # of the blended vectors
    m = (0.6 - 0.25) / (1.0 - 0.25)
    a, _ = O.iterate(f, v, j, rout=0.6)
    pb, _ = O.iterate(bulb, v, j)
    px, _ = O.iterate(box, v, j)
    mix = m * px[:3] + (1 - m) * pb[:3]
    length = m * np.linalg.norm(px[:3]) + (1 - m) * np.linalg.norm(pb[:3])
    assert abs(np.linalg.norm(a[:3]) - length) < 1e-12
    assert np.abs(a[:3] / np.linalg.norm(a[:3]) - mix / np.linalg.norm(mix)).max() < 1e-12


def test_bulbox_bulb_scaling():
    """int_pow2_scale(s) = s * P2(v / s) + C."""
    f = H.formula("Bulbox", bulb_scale=1.7, r_threshold=1.0, r_threshold2=1.0)
    v = np.array([0.3, -0.2, 0.25, 0.0])
    j = [0.05, -0.1, 0.2, 0.0]
    a, _ = O.iterate(f, v, j, rout=0.2)
    b, _ = O.iterate(H.formula("Integer Power", power=2, z_mul=-1.0), v / 1.7, [0.0] * 4)
    assert np.abs(a[:3] - (1.7 * b[:3] + j[:3])).max() < 1e-12


def test_folding_int_pow():
    """The fold leaves |x| <= fold alone (x = |x+f| - |x-f| - x) and mirrors
    x > f to 2f - x, then the integer power."""
    for power in (2, 5, 8):
        f = H.formula("Folding Int Pow", power=float(power), z_mul=-1.0, fold=2.0)
        for p in POINTS:
            a, _ = O.iterate(f, [p[0], p[1], p[2], 0.0], [0.1, 0.0, -0.1, 0.0])
            b, _ = O.iterate(H.formula("Integer Power", power=power), [p[0], p[1], p[2], 0.0], [0.1, 0.0, -0.1, 0.0])
            assert np.abs(a[:3] - b[:3]).max() < 1e-12
    f = H.formula("Folding Int Pow", power=2.0, z_mul=-1.0, fold=0.5)
    a, _ = O.iterate(f, [0.8, -0.9, 0.2, 0.0], [0.0] * 4)
    b, _ = O.iterate(H.formula("Integer Power", power=2), [1.0 - 0.8, -1.0 + 0.9, 0.2, 0.0], [0.0] * 4)
    assert np.abs(a[:3] - b[:3]).max() < 1e-12


def test_smatrix4_by_hand_and_orthonormal():
    s = H.scene_preset("Quaternion")
    s.rot_4d[:] = [0.7, 0.0, 0.0]
    m = np.array(O.params(s).smatrix4)
    c, sn = float(np.float32(math.cos(0.7))), float(np.float32(math.sin(0.7)))
    assert m[0, 0] == c and m[3, 3] == c and m[0, 3] == -sn and m[3, 0] == sn
    assert m[1, 1] == 1.0 and m[2, 2] == 1.0
    s.rot_4d[:] = [0.3, -0.9, 1.4]
    m = np.array(O.params(s).smatrix4)
    assert np.abs(m @ m.T - np.eye(4)).max() < 1e-6
    assert abs(np.linalg.det(m) - 1.0) < 1e-6
    # the XW rotation comes first: a rotation of XW alone then YW alone
    rx = np.eye(4)
    rx[0, 0] = rx[3, 3] = math.cos(0.3)
    rx[0, 3], rx[3, 0] = -math.sin(0.3), math.sin(0.3)
    ry = np.eye(4)
    ry[1, 1] = ry[3, 3] = math.cos(-0.9)
    ry[1, 3], ry[3, 1] = -math.sin(-0.9), math.sin(-0.9)
    rz = np.eye(4)
    rz[2, 2] = rz[3, 3] = math.cos(1.4)
    rz[2, 3], rz[3, 2] = -math.sin(1.4), math.sin(1.4)
    assert np.abs(m - rz @ ry @ rx).max() < 1e-6


def test_julia_mode_uses_the_seed():
    """In Julia mode the constant is the seed, not the point: a 4D Julia set
    at seed (a, b, 0, 0) restricted to the complex plane is the complex Julia
    set of a + bi."""
    s = H.scene_preset("Quaternion")
    s.iterations = 40
    s.julia_c[:] = [-0.4, 0.6, 0.0, 0.0]
    p = O.params(s)
    rng = np.random.default_rng(5)
    for _ in range(40):
        x, y = rng.uniform(-1.5, 1.5), rng.uniform(-1.2, 1.2)
        _, its = O.de_at(p, [x, y, 0.0])
        z, c, want = complex(x, y), complex(-0.4, 0.6), 40
        for it in range(1, 41):
            z = z * z + c
            if abs(z) ** 2 > p.d_rstop:
                want = it
                break
        assert its == want, (x, y)


def ipol_scene(f0, f1, w, name="Integer Power", size=(96, 72)):
    s = H.scene_preset(name, *size)
    H.set_formulas(s, [f0, f1])
    s.interpolation = 1
    s.interpolation_w[:] = w
    return s


def test_interpolation_weights_are_normalised():
    """GetMCTparasFromHeader's normalisation (src/calc.rs:378-397) and the
    interpolation DE scale: the weighted sum of the formulas' scales."""
    p = O.params(ipol_scene(H.formula("Integer Power"), H.formula("Bulbox"), [3.0, 1.0]))
    assert p.ipol and list(p.ipol_w) == [np.float32(0.75), np.float32(0.25)]
    de_offset = np.float32(min(np.float32(p.de_stop) * np.float32(0.1), np.float32(0.004)))
    assert p.d_de_scale == np.float32(np.float32(0.75 * 1.0 + 0.25 * 0.2) * de_offset)
    # no positive weight: no formula gives the DE option (an error, as the reference's)
    with pytest.raises(RuntimeError, match="not valid"):
        O.params(ipol_scene(H.formula("Integer Power"), H.formula("Bulbox"), [0.0, 0.0]))
    # a weight <= 0 leaves its formula's DE option out: the box alone is analytic
    p = O.params(ipol_scene(H.formula("Integer Power"), H.formula("Amazing Box"), [-1.0, 1.0]))
    assert p.de_option == 11 and p.is_custom_de
    p = O.params(ipol_scene(H.formula("Quaternion"), H.formula("Quaternion", yw_mul=0.5), [1.0, 1.0]))
    assert p.de_option == 4 and p.four_d


def test_interpolation_of_a_formula_with_itself_is_that_formula():
    """Blending two equal results gives the result: the picture of
    (F, F) at any weights is that of F alone, up to the last bit of the
    blended radius."""
    f = H.formula("Integer Power", power=5)
    a = O.render(O.params(ipol_scene(f, f, [0.3, 0.7])), threads=4)
    s = H.scene_preset("Integer Power", 96, 72)
    H.set_formulas(s, [f])
    b = O.render(O.params(s), threads=4)
    both = np.isfinite(a["depth"]) & np.isfinite(b["depth"])
    assert (np.isfinite(a["depth"]) == np.isfinite(b["depth"])).mean() > 0.998
    assert np.abs(a["depth"][both] - b["depth"][both]).max() < 1e-3
    assert (a["iters"] == b["iters"]).mean() > 0.998


def test_interpolation_is_symmetric():
    """(F0, F1, w0, w1) = (F1, F0, w1, w0): the blend is a sum of two
    products, whatever the order."""
    f0, f1 = H.formula("Integer Power", power=3), H.formula("Tricorn")
    a = O.render(O.params(ipol_scene(f0, f1, [0.6, 0.4])), threads=4)
    b = O.render(O.params(ipol_scene(f1, f0, [0.4, 0.6])), threads=4)
    assert np.array_equal(a["depth"], b["depth"]) and np.array_equal(a["iters"], b["iters"])
