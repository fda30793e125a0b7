"""The f64 CPU oracle pinned against the reference's own known-answer tests
and against what the mathematics fixes (CPU only).

Reference tests re-expressed here:
  src/formulas.rs:864-878  a negative Min R acts as 0
  src/formulas.rs:915-942  Integer Power 2..8 = the spherical triplex formula (1e-9)
  src/formulas.rs:957-966  the box scales the derivative by scale / MinR^2
  src/calc.rs:2985-2997    pixel_seed's jump-ahead = stepping the generator
  tests/render.rs:12-22    the G-buffer does not depend on the thread count
  tests/render.rs:24-31    the presets' coverage lies in (0.05, 0.995) at 96 x 72
  tests/render.rs:47-61    bulb DE: origin at max iterations, (0.3, 0.2, -2) in (0.2, 1.5)
  tests/render.rs:63-72    box analytic DE at (0, 0, -9) in (0.5, 6)
  tests/render.rs:118-140  hard shadows: some pixel in shadow
  tests/render.rs:162-178  DE AO: some pixel with amb_shadow > 2000
  src/deao.rs:464-470      DE AO ray counts
"""
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O


def preset(name, w=96, h=72):
    """Scene::preset (src/scene.rs:811-838) at the test size."""
    return H.scene_preset(name, w, h)


def triplex(n, v):
    r = math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    th = math.atan2(v[1], v[0]) * n
    ph = math.atan2(v[2], math.hypot(v[0], v[1])) * n
    rn = r ** n
    return [rn * math.cos(th) * math.cos(ph), rn * math.sin(th) * math.cos(ph), -rn * math.sin(ph)]


@pytest.mark.parametrize("n", range(2, 9))
def test_int_pow_matches_spherical_triplex(n):
    for p in ([0.31, -0.42, 0.27], [-0.7, 0.1, 0.45], [0.05, 0.9, -0.3]):
        v, _ = O.iterate(H.intpow(n, 1), [p[0], p[1], p[2], 0.0], [0.0] * 4)
        assert np.allclose(v[:3], triplex(n, p), atol=1e-9, rtol=0)


def test_int_pow_adds_the_constant():
    c = [0.12, -0.3, 0.05, 0.0]
    p = [0.31, -0.42, 0.27, 0.0]
    v, _ = O.iterate(H.intpow(8, 1), p, c)
    t = triplex(8, p)
    assert np.allclose(v[:3], [t[k] + c[k] for k in range(3)], atol=1e-12)


def test_box_scales_derivative_inside_min_r():
    v, _ = O.iterate(H.amazing_box(2.0, 0.5, 1.0), [0.1, 0.1, 0.1, 1.0], [0.0] * 4, ade=True)
    assert abs(v[3] - 8.0) < 1e-12


def test_box_fold_and_spheres_by_hand():
    # fold: x -> |x+f| - |x-f| - x; outside the unit sphere the scale alone
    v, _ = O.iterate(H.amazing_box(2.0, 0.5, 1.0), [1.5, 0.2, -3.0, 1.0], [0.0] * 4, ade=True)
    folded = [0.5, 0.2, 1.0]  # 1.5 -> 2.5 - 0.5 - 1.5, -3 -> 2 - 4 + 3
    assert np.allclose(v[:3], [2.0 * x for x in folded]) and abs(v[3] - 2.0) < 1e-15
    # between min R^2 and 1: scale / r^2
    v, _ = O.iterate(H.amazing_box(2.0, 0.5, 1.0), [0.6, 0.0, 0.0, 1.0], [0.0] * 4, ade=True)
    assert abs(v[0] - 0.6 * 2.0 / 0.36) < 1e-12 and abs(v[3] - 2.0 / 0.36) < 1e-12


def test_negative_min_r_acts_as_zero():
    for ade in (False, True):
        for p in ([0.1, 0.1, 0.1, 1.0], [0.8, -0.3, 0.2, 1.0], [0.01, 0.02, -0.01, 1.0]):
            neg, _ = O.iterate(H.amazing_box(1.7, -0.6, 0.3), p, [0.0] * 4, ade)
            zero, _ = O.iterate(H.amazing_box(1.7, 0.0, 0.3), p, [0.0] * 4, ade)
            pos, _ = O.iterate(H.amazing_box(1.7, 0.6, 0.3), p, [0.0] * 4, ade)
            assert np.array_equal(neg, zero)
            assert not np.array_equal(neg, pos)


def test_pixel_seed_steps_along_the_row():
    for y in (0, 7, 1499):
        s = np.int64((0x24563487 + (y + 1) * 0x324594A1) & 0xFFFFFFFF)
        for x in range(3000):
            got = O.pixel_seed(y, x) & 0xFFFFFFFF
            assert got == int(s), (y, x)
            s = (s * 214013 + 2531011) & 0xFFFFFFFF


def test_bulb_distance_estimate_is_sane():
    p = O.params(preset("Integer Power"))
    _, its = O.de_at(p, [0.0, 0.0, 0.0])
    assert its == 12
    d, its = O.de_at(p, [0.3, 0.2, -2.0])
    assert its < 12 and 0.2 < d < 1.5


def test_box_analytic_distance_estimate_is_sane():
    p = O.params(preset("Amazing Box"))
    assert p.is_custom_de and p.de_option == 11
    d, _ = O.de_at(p, [0.0, 0.0, -9.0])
    assert 0.5 < d < 6.0


@pytest.mark.parametrize("name", sorted(H.FORMULA_NAMES))
def test_presets_hit_the_object(name):
    r = O.render(O.params(preset(name)), threads=4)
    obj = np.isfinite(r["depth"])
    assert 0.05 < obj.mean() < 0.995


def test_gbuffer_is_thread_independent():
    p = O.params(preset("Integer Power"))  # first step random on: the pixel seeds
    a = O.render(p, threads=1)
    b = O.render(p, threads=3)
    assert np.array_equal(a["gbuf"], b["gbuf"])
    assert np.array_equal(a["depth"], b["depth"])


def test_analytic_and_numeric_de_agree_in_sign_and_magnitude():
    sa = H.scene_bulb(96, 72, seed=2, de_mode=H.DE_ANALYTIC_P8)
    sn = H.scene_bulb(96, 72, seed=2, de_mode=H.DE_REFERENCE)
    pa, pn = O.params(sa), O.params(sn)
    rng = np.random.default_rng(5)
    ratios = []
    for _ in range(200):
        q = rng.normal(size=3)
        q = q / np.linalg.norm(q) * rng.uniform(1.1, 2.0)
        da, ia = O.de_at(pa, q)
        dn, _ = O.de_at(pn, q)
        if ia < 60 and da > 1e-4:
            ratios.append(dn / da)
    ratios = np.array(ratios)
    assert (ratios > 0).all()
    assert 0.1 < np.median(ratios) < 10.0


def test_rays_that_miss_the_bounding_sphere_are_background():
    sc = H.scene_bulb(64, 48, seed=4)
    p = O.params(sc)
    r = O.render(p, threads=4)
    for y in range(0, 48, 3):
        for x in range(0, 64, 3):
            c, v = O.pixel_ray(p, x, y)
            u = v / np.linalg.norm(v)
            t = -np.dot(c, u)
            closest = np.linalg.norm(c + max(t, 0.0) * u)
            if closest > 1.6:
                assert not np.isfinite(r["depth"][y * 64 + x]), (x, y)


def test_march_stops_at_the_first_de_crossing():
    """Brute force along the ray: the first point where the DE falls under
    the DE stop lies within the last step before the surface found."""
    sc = H.scene_bulb(48, 36, seed=6, de_mode=H.DE_ANALYTIC_P8)
    sc.bin_search_steps = 30
    p = O.params(sc)
    r = O.render(p, threads=4)
    checked = agree = 0
    for i in np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))[::7]:
        x, y = i % 48, i // 48
        c, v = O.pixel_ray(p, x, y)
        t_hit = r["depth"][i]
        ts = np.arange(0.0, t_hit + 2.0, 0.02)
        first = None
        for t in ts:
            d, its = O.de_at(p, c + t * v)
            ms = p.de_stop * (1.0 + t * p.de_stop_factor)
            if its >= p.max_it or d / p.step_width < ms:
                first = t
                break
        checked += 1
        if first is not None and abs(first - t_hit) < 0.05:
            agree += 1
    assert checked > 20 and agree / checked > 0.9, (agree, checked)


def test_normals_follow_the_de_gradient():
    """The normal stencil (src/calc.rs:1850-1858) against central
    differences of the DE at a smaller step, in the view frame."""
    sc = H.scene_bulb(48, 36, seed=7, de_mode=H.DE_ANALYTIC_P8)
    p = O.params(sc)
    r = O.render(p, threads=4)
    vg = np.array([[p.vgrads[i][k] for k in range(3)] for i in range(3)])
    cosines = []
    for i in np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))[::5]:
        x, y = i % 48, i // 48
        c, v = O.pixel_ray(p, x, y)
        q = c + r["depth"][i] * v
        h = 0.02
        g = np.zeros(3)
        for k in range(3):
            e = vg[k] * h
            g[k] = (O.de_at(p, q + e)[0] - O.de_at(p, q - e)[0])
        g /= np.linalg.norm(g)
        cosines.append(float(np.dot(g, r["normals"][i])))
    cosines = np.array(cosines)
    assert np.median(cosines) > 0.99 and (cosines > 0.9).mean() > 0.9


def test_depth_code_round_trip():
    """calc_zpos_and_rough (src/calc.rs:2194-2205) and its inverse
    (src/calc.rs:2167): the code's middle lies within one digit."""
    sc = H.scene_bulb(64, 48, seed=8)
    p = O.params(sc)
    r = O.render(p, threads=4)
    hit = np.isfinite(r["depth"]) & (r["depth"] > 0)
    zf = (r["gbuf"][hit, 2] >> 8).astype(np.float64)
    a = (8388351.5 - zf) / p.zc_mul + 1.0
    back = (a * a - 1.0) / p.zcorr
    dz = p.zc_mul * p.zcorr / (2.0 * np.sqrt(r["depth"][hit] * p.zcorr + 1.0))  # code per step width
    assert (np.abs(back - r["depth"][hit]) * dz < 1.0).all()


def test_shadows_mark_pixels():
    sc = preset("Integer Power")
    sc.first_step_random = 0
    p = O.params(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], params=p)
    r = O.render(p, job, threads=4)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    shadowed = (r["gbuf"][:, 1] >> 16) & 0x400
    assert surf.any() and (shadowed[surf] != 0).any()
    assert (shadowed[~surf] == 0).all()
    # a light from straight behind the camera lights more than one from the side
    back = O.render(p, H.post_job(lights=[(0.0, 0.0, -1.0)], params=p), threads=4)
    sb = (back["gbuf"][:, 1] >> 16) & 0x400
    assert (sb[surf] != 0).mean() < (shadowed[surf] != 0).mean()


@pytest.mark.parametrize("preset,light", [("Integer Power", (0.6, -0.5, -0.6)), ("Amazing Box", (-0.3, 0.8, -0.5)),
                                          ("Integer Power", (0.0, 0.9, 0.4))])
def test_hard_shadows_against_brute_force(preset, light):
    """hard_shadow_row's bit against a brute-force walk towards the light:
    from the surface point 0.1 step back along the view ray, the DE sampled
    every 0.05 step over the pass's reach (hs_max_len, src/calc.rs:2767-2775);
    blocked where it reaches the DE stop of its depth or the iteration limit,
    shadowed also where the surface faces away.  The pass never shadows a
    pixel the walk finds lit; it may miss an occluder thinner than its step."""
    from test_oracle_mc import view_tasks, world_normals

    sc = H.scene_preset(preset, 48, 36)
    sc.first_step_random = 0
    p = O.params(sc)
    job = H.post_job(lights=[light], params=p)
    r = O.render(p, job, threads=0)
    t = view_tasks(p)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    bit = ((r["gbuf"][surf, 1] >> 16) & 0x400) != 0
    lvec = np.array(job.light_vec[0][:])
    vf = t["dir"][surf]
    mzz = np.maximum(r["depth"][surf] - 0.1, 0.0)
    start = t["pos"][surf] + vf * (r["depth"][surf] - 0.1)[:, None]
    zz2mul = -(vf @ lvec) / np.sqrt((vf * vf).sum(1) * (lvec @ lvec))
    y = surf // p.width
    f32 = np.float32
    mz = np.minimum(mzz, p.zend * 0.4).astype(f32)
    reach = ((p.width + y).astype(f32) * f32(0.6) * (f32(1) + f32(0.5) * mz * f32(max(p.fov_y, 0)) / f32(p.height))
             * f32(job.shadow_max_len_mul)).astype(np.float64)
    away = world_normals(p, r["normals"])[surf] @ lvec > 0
    blocked = np.zeros(len(surf), bool)
    for i in range(len(surf)):
        s = np.arange(0.05, reach[i] + 1e-9, 0.05)
        de, its = O.de_many(p, start[i] + np.outer(s, -lvec))
        ms = p.de_stop * (1.0 + np.abs(mzz[i] + s * zz2mul[i]) * p.de_stop_factor)
        blocked[i] = np.any((de <= ms) | (its >= p.max_it))
    brute = blocked | away
    assert 0.05 < brute.mean() < 0.95
    assert not np.any(bit & ~brute)
    assert (bit == brute).mean() >= 0.99, ((bit == brute).mean(), bit.mean(), brute.mean())


def test_deao_occludes():
    sc = preset("Integer Power")
    sc.first_step_random = 0
    p = O.params(sc)
    for q in (1, 4):
        r = O.render(p, H.post_job(deao_quality=q), threads=4)
        surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
        amb = r["gbuf"][:, 4] & 0xFFFF
        assert (amb[surf] > 2000).any()
        assert (amb[surf] < 16383).all()
