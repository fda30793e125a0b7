"""The oracle's general surface modes (test infrastructure, CPU):
calculate_normals with smooth normals 1..8 and the roughness
(src/calc.rs:1806-1903), calculate_normals_on_smooth_it (1910-2003),
do_color_on_it and do_color's options 1..5 (2010-2090), pinned by what the
mathematics fixes:
  - create_xy_vecs_from_normals gives two unit vectors across the normal
    and across each other (src/math.rs:218-232);
  - the smoothing sweeps move the stencil's normal towards the gradient of
    the DE over a wider neighbourhood, and the roughness is zero where the
    DE is linear across the sweeps (a flat field: dsg = 0);
  - the normal on the smooth iteration count points the way of the normal
    on the DE (the count falls where the DE rises);
  - the reference's own scene fixture, tests/data/mandelbulb.m3s (smooth
    normals 3), renders with the reference's coverage window and a
    roughness byte.
Colour options 1..5 and colour on an iteration: the orbit end they read is
pinned by an independent iterate of the surface point (color_on_it: the
point after color_on_it - 1 formula steps), and each word inverts to that
orbit end -- the angle words' cosine and sine against the normalised
projections of v - c (options 2..4) or of v (option 5, azimuth and
elevation), option 1's logarithm back to the radii -- so a swapped atan2
operand, a wrong offset or scale, or a wrong orbit point fails.  The maps
have no known answer beyond their definitions (the reference's own test
checks only that colour on an iteration changes the colouring)."""
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O


def test_xy_vecs_are_orthonormal_across_the_normal():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = rng.normal(size=3)
        vx, vy = O.xy_vecs(n)
        u = n / np.linalg.norm(n)
        assert abs(np.linalg.norm(vx) - 1) < 1e-12 and abs(np.linalg.norm(vy) - 1) < 1e-12
        assert abs(vx @ vy) < 1e-12 and abs(vx @ u) < 1e-12 and abs(vy @ u) < 1e-12
    vx, vy = O.xy_vecs([0.0, 0.0, -1.0])  # along the view axis: the screen axes
    assert np.allclose(vx, [1, 0, 0]) and np.allclose(vy, [0, 1, 0])


def world_normal(p, nv):
    vg = np.array([[p.vgrads[i][k] for k in range(3)] for i in range(3)])
    w = nv @ (vg / np.linalg.norm(vg, axis=1, keepdims=True))
    return w / np.linalg.norm(w)


def test_smoothing_moves_the_normal_towards_the_wider_gradient():
    sc0 = H.scene_bulb(96, 72, seed=5, de_mode=H.DE_ANALYTIC_P8)
    sc3 = H.scene_bulb(96, 72, seed=5, de_mode=H.DE_ANALYTIC_P8)
    sc3.smooth_normals = 3
    p0, p3 = O.params(sc0), O.params(sc3)
    r0, r3 = O.render(p0, threads=0), O.render(p3, threads=0)
    assert np.array_equal(r0["depth"], r3["depth"])  # the march does not change
    vg = np.array([[p0.vgrads[i][k] for k in range(3)] for i in range(3)])
    surf = np.flatnonzero(np.isfinite(r0["depth"]) & (r0["depth"] > 0))[::3]
    a0, a3 = [], []
    for i in surf:
        c, v = O.pixel_ray(p0, i % 96, i // 96)
        q = c + r0["depth"][i] * v
        h = 0.6 * p0.de_stop  # the sweeps reach 2 x 0.15 x 3 steps out: a wider difference
        g = np.array([O.de_at(p0, q + vg[k] * h)[0] - O.de_at(p0, q - vg[k] * h)[0] for k in range(3)])
        g /= np.linalg.norm(g)
        a0.append(math.degrees(math.acos(np.clip(g @ r0["normals"][i], -1, 1))))
        a3.append(math.degrees(math.acos(np.clip(g @ r3["normals"][i], -1, 1))))
    assert np.mean(a3) < np.mean(a0), (np.mean(a3), np.mean(a0))
    # the roughness byte: set with smooth normals only, and some pixels rough
    rough0 = r0["gbuf"][surf, 2] & 0xFF
    rough3 = r3["gbuf"][surf, 2] & 0xFF
    assert not rough0.any() and rough3.any() and (rough3 < 255).mean() > 0.5


@pytest.mark.parametrize("sm", [1, 8])
def test_other_smoothings_stay_near_the_stencil(sm):
    sc = H.scene_bulb(64, 48, seed=6, de_mode=H.DE_ANALYTIC_P8)
    p0 = O.params(sc)
    sc.smooth_normals = sm
    p = O.params(sc)
    r0, r = O.render(p0, threads=0), O.render(p, threads=0)
    surf = np.isfinite(r0["depth"]) & (r0["depth"] > 0)
    cosang = (r0["normals"][surf] * r["normals"][surf]).sum(1)
    assert np.median(cosang) > math.cos(math.radians(15)), np.median(cosang)


def test_normals_on_the_smooth_iteration_point_like_the_de_normals():
    sc = H.scene_preset("Integer Power", 96, 72)
    sc.first_step_random = 0
    p_de = O.params(sc)
    sc.normals_on_de = 0
    p_it = O.params(sc)
    assert not p_it.normals_on_de
    a, b = O.render(p_de, threads=0), O.render(p_it, threads=0)
    surf = np.isfinite(a["depth"]) & (a["depth"] > 0)
    cosang = (a["normals"][surf] * b["normals"][surf]).sum(1)
    assert np.median(cosang) > math.cos(math.radians(20)), np.median(cosang)
    assert (cosang > 0).mean() > 0.95  # a flipped sign would turn them around


def test_the_reference_fixture_renders():
    """tests/data/mandelbulb.m3s at 200 x 150: the coverage window of the
    reference's render tests (tests/render.rs:24-31), the roughness byte of
    smooth normals 3."""
    sc = H.scene_m3s_mandelbulb(200, 150)
    p = O.params(sc)
    assert p.sm_normals == 3 and not p.is_custom_de and p.de_add_steps == 8
    r = O.render(p, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    assert 0.05 < surf.mean() < 0.995, surf.mean()
    assert (r["gbuf"][surf, 2] & 0xFF).any()


@pytest.mark.parametrize("option", [1, 2, 3, 4, 5])
def test_colour_options_fill_their_words(option):
    sc = H.scene_bulb(64, 48, seed=7, de_mode=H.DE_ANALYTIC_P8)
    sc.color_option = option
    p = O.params(sc)
    r = O.render(p, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    otrap = r["gbuf"][surf, 3] >> 16
    assert otrap.max() <= 32767 and len(np.unique(otrap)) > 20
    if option in (2, 3, 4):  # (atan2 + pi) 5200: within [0, 2 pi 5200]
        assert otrap.max() <= round(2 * math.pi * 5200)


def test_colour_on_the_start_point_differs():
    sc = H.scene_bulb(64, 48, seed=7, de_mode=H.DE_ANALYTIC_P8)
    sc.color_option = 2
    a = O.render(O.params(sc), threads=0)
    sc.color_on_it = 1  # the orbit "end" is the point itself: atan2 of zero vectors
    b = O.render(O.params(sc), threads=0)
    surf = np.isfinite(a["depth"]) & (a["depth"] > 0)
    assert np.array_equal(a["depth"], b["depth"])
    ob = b["gbuf"][surf, 3] >> 16
    assert (ob == round(math.pi * 5200)).all()  # atan2(0, 0) + pi


def surface_probes(sc, n=40, seed=0):
    p = O.params(sc)
    r = O.render(p, threads=0)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    rng = np.random.default_rng(seed)
    out = []
    for i in rng.choice(surf, size=min(n, surf.size), replace=False):
        pr = O.color_probe(p, int(i % p.width), int(i // p.width))
        assert pr is not None
        # the probe's march is the render's: the same words at that pixel
        assert pr["otrap"] == r["gbuf"][i, 3] >> 16, i
        out.append((pr, r["gbuf"][i]))
    return p, out


@pytest.mark.parametrize("option", [1, 2, 3, 4, 5])
def test_colour_words_invert_to_the_orbit_end(option):
    sc = H.scene_bulb(64, 48, seed=7, de_mode=H.DE_ANALYTIC_P8)
    sc.color_option = option
    p, probes = surface_probes(sc)
    for pr, g in probes:
        v, c = pr["v"], pr["c"]
        assert abs(v @ v - pr["rout"]) <= 1e-12 * pr["rout"]  # rout: the orbit end's squared radius
        w = pr["otrap"]
        assert 0 <= w <= 32767
        if option == 1:
            mul = float(np.float32(sc.color_mul * 512.0))
            want = math.log(pr["rout"] / (pr["rold"] + 1.0))
            if 0 < w < 32767:
                assert abs(w / mul - want) <= 1.0 / mul
        elif option in (2, 3, 4):
            i, j = {2: (0, 1), 3: (0, 2), 4: (1, 2)}[option]
            d = np.array([v[i] - c[i], v[j] - c[j]])
            a = w / 5200.0 - math.pi
            assert np.linalg.norm(np.array([math.cos(a), math.sin(a)]) - d / np.linalg.norm(d)) <= 2.0 / 5200.0
        else:
            rho = math.hypot(v[0], v[1])
            a = w / 5215.0 - math.pi  # the azimuth from the y axis: atan2(x, y)
            assert np.linalg.norm(np.array([math.sin(a), math.cos(a)]) - v[:2] / rho) <= 2.0 / 5215.0
            e = (pr["gradient"] / 5215.0 - math.pi) / 2.0  # the elevation: asin(z / r)
            assert abs(math.sin(e) - v[2] / np.linalg.norm(v)) <= 2.0 / 5215.0


@pytest.mark.parametrize("k", [2, 3, 5])
def test_colour_on_an_iteration_is_the_orbit_after_k_minus_1_steps(k):
    """colour_on_it = k: the orbit point after k - 1 formula steps from the
    surface point (Formula::iterate, pinned by the triplex closed form in
    test_oracle.py), unless it escapes (rstop3d) first."""
    sc = H.scene_bulb(64, 48, seed=7, de_mode=H.DE_ANALYTIC_P8)
    sc.color_option = 2
    sc.color_on_it = k
    p, probes = surface_probes(sc, n=25, seed=k)
    f = H.intpow(8, 1)
    for pr, _ in probes:
        c = pr["c"]
        v = np.array([c[0], c[1], c[2], 0.0])
        j = np.array([c[0], c[1], c[2], 0.0])
        for _ in range(k - 1):
            v, j = O.iterate(f, v, j, ade=True)
            if v[:3] @ v[:3] > p.rstop3d:
                break
        assert np.array_equal(v[:3], pr["v"]), (k, v, pr["v"])


def test_colour_on_iteration_2_changes_the_orbit_trap_colouring():
    """The reference's own check (tests/render.rs:233-241): colour on
    iteration 2 (color_on_it 3) changes the orbit-trap colouring of the
    picture, and leaves the depths."""
    sc = H.scene_bulb(64, 48, seed=7, de_mode=H.DE_ANALYTIC_P8)
    a = O.render(O.params(sc), threads=0)
    sc.color_on_it = 3
    b = O.render(O.params(sc), threads=0)
    surf = np.isfinite(a["depth"]) & (a["depth"] > 0)
    assert np.array_equal(a["depth"], b["depth"])
    ot = lambda r: int((r["gbuf"][surf, 3] >> 16).astype(np.uint64).sum())
    assert ot(a) != ot(b)
