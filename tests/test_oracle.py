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


def _sphere_dirs(n):
    """n directions spread evenly over the sphere (a Fibonacci lattice)."""
    i = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    th = np.pi * (1.0 + 5.0 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


def _inside(p, pts):
    return O.de_many(p, pts)[1] >= p.max_it


def _distance_to_set(p, q, dirs, reach, samples=160, halvings=16):
    """An upper bound of the distance from q to the set (the points whose
    orbit does not escape within max_it iterations, decided by the
    iteration count alone): along each direction the first sample inside,
    bisected; the nearest of the directions."""
    ts = np.linspace(0.0, reach, samples + 1)[1:]
    ins = _inside(p, (q[None, None, :] + dirs[:, None, :] * ts[None, :, None]).reshape(-1, 3)).reshape(len(dirs), -1)
    hit = ins.any(1)
    if not hit.any():
        return np.inf
    k = np.flatnonzero(hit)
    j = np.argmax(ins[k], 1)
    lo = np.where(j > 0, ts[np.maximum(j - 1, 0)], 0.0)
    hi = ts[j]
    for _ in range(halvings):
        m = 0.5 * (lo + hi)
        inn = _inside(p, q[None, :] + dirs[k] * m[:, None])
        hi = np.where(inn, m, hi)
        lo = np.where(inn, lo, m)
    return float(hi.min())


def _de_over_distance(de_mode, n_points=24, seed=3):
    """DE / brute-force distance at points outside the bulb: each point a
    random distance (0.02 to 0.25) beyond the set's boundary along a random
    ray from the origin (bisected on the iteration count)."""
    p = O.params(H.scene_bulb(96, 72, seed=2, de_mode=de_mode))
    rng = np.random.default_rng(seed)
    dirs = _sphere_dirs(300)
    ratios = []
    while len(ratios) < n_points:
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        lo, hi = 0.0, 2.0
        for _ in range(40):
            m = 0.5 * (lo + hi)
            if _inside(p, (u * m)[None, :])[0]:
                lo = m
            else:
                hi = m
        q = u * (hi + rng.uniform(0.02, 0.25))
        d, its = O.de_at(p, q)
        dist = _distance_to_set(p, q, dirs, reach=1.2)
        if its >= p.max_it or not np.isfinite(dist) or dist < 1e-3:
            continue  # on a filament of the set: no distance to measure
        ratios.append(d / dist)
    return np.array(ratios)


def test_analytic_p8_de_is_a_distance_bound_of_the_right_size():
    """The analytic power-8 DE (0.5 r ln r / dr, src/iteration.rs:562-568,
    in world units) against the distance to the set found by brute force
    (the first non-escaping point along 300 directions, src/calc.rs:1515-1545
    marches to where that distance vanishes): a lower bound for nearly all
    points, and within a factor of four.  Twice the constant puts the upper
    tail above the distance, half of it the median under a quarter."""
    r = _de_over_distance(H.DE_ANALYTIC_P8)
    assert np.percentile(r, 90) <= 1.0, np.sort(r)
    assert 0.25 <= np.median(r) <= 0.65, np.sort(r)


def test_numeric_de_has_the_size_of_the_distance():
    """The reference's numeric gradient DE (option 0, src/calc.rs:1515-1545:
    rout ln rout over the finite-difference gradient, twice the analytic
    bound) against the same brute-force distance: the median near the
    distance itself.  A factor of two either way leaves the window."""
    r = _de_over_distance(H.DE_REFERENCE)
    assert 0.6 <= np.median(r) <= 1.3, np.sort(r)
    assert np.percentile(r, 90) <= 1.6, np.sort(r)


def _rodrigues(a, b):
    """The rotation taking unit vector a to unit vector b about the axis
    a x b (the smallest such rotation)."""
    k = np.cross(a, b)
    s, c = np.linalg.norm(k), float(np.dot(a, b))
    if s < 1e-12:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    k /= s
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + s * K + (1.0 - c) * K @ K


def _ao_set(quality, dt1=None, dt2=0.0):
    """The DE AO ray set of a quality in the local frame of a surface whose
    normal is (0, 0, -1) (the view looks along +z): the straight ray, then
    rings at polar angles y abr (abr = (pi / 2) / (quality + 0.9)), each
    with round(sin(y abr) 2 pi / abr) rays evenly spaced in azimuth from the
    +x side (src/deao.rs:75-105); quality 0: three rays at 30 degrees.
    Dithered (dt1, dt2 = the pixel's place in its cell times 1/2 / dither):
    the rings at (y + dt1 - 1/4) abr and azimuths shifted by dt2 of a gap,
    the straight ray dropped when dt1 > 0.1; quality 0 at (dt1 + 1/2) 50
    degrees (src/deao.rs:318-342)."""
    out = []
    if quality == 0:
        polar = np.pi / 6.0 if dt1 is None else (dt1 + 0.5) * np.radians(50.0)
        rings, abr = [(polar, 3)], None
    else:
        abr = (np.pi / 2.0) / (quality + 0.9)
        if dt1 is None or dt1 <= 0.1:
            out.append((0.0, 0.0))
        shift = 0.0 if dt1 is None else dt1 - 0.25
        rings = [((y + shift) * abr, int(round(np.sin(y * abr) * 2 * np.pi / abr))) for y in range(1, quality + 1)]
    for polar, n in rings:
        out += [(polar, 2 * np.pi * (x + dt2) / n) for x in range(n)]
    return np.array([[np.sin(a) * np.cos(z), -np.sin(a) * np.sin(z), -np.cos(a)] for a, z in out])


def _object(g):
    """The G-buffer records that are surface pixels (zpos and gradient
    words below 32768)."""
    return ((g[:, 2] >> 16) < 32768) & ((g[:, 3] & 0xFFFF) < 32768)


def _row_state(p, g, x, y, draws):
    """The DE AO generator's state before pixel (x, y): the row's seed
    advanced by `draws` numbers for each surface pixel left of it."""
    before = int(_object(g[y * p.width: y * p.width + x]).sum())
    seed = O.deao_row_seed(y)
    if before == 0:
        return seed
    return int(O.deao_rand(seed, before * draws)[0][-1])


@pytest.mark.parametrize("quality,dither,rnd", [(1, 0, False), (2, 0, False), (1, 1, True), (2, 2, False),
                                               (0, 1, True), (3, 2, True)])
def test_deao_against_brute_force(quality, dither, rnd):
    """deao_row (src/deao.rs:268-418) per pixel, on a sample of surface
    pixels, against what it must compute:
      - the rebuilt surface point (surface_point, src/calc.rs:2157-2192)
        lies on the pixel's ray at the marched depth;
      - each ray's world direction is the ray set (rings around the straight
        ray) turned by the smallest rotation that takes the view axis's
        reverse to the pixel's normal, then into the world frame by the
        view's axes (Rodrigues, independent of the quaternion of
        normal_matrix);
      - each ray's openness min over its samples of (DE - stop + t md) / t
        (capped at 1) is never below the same minimum over a fine walk of
        the DE along the ray (0.05 steps), and close to it for most rays
        (the pass's samples are sparser: median gap 0.006-0.009 here; a
        dropped DE stop makes it 0.06);
      - a pixel with every ray open has amb 0.
    Dithered, the set of the pixel's cell; with the random first step, each
    ray's first sample at 0.5..2 of the plain first step, the walk starting
    there (not held to the ray's length).
    A swapped axis of the AO frame, a sign of its quaternion, a dropped DE
    stop or swapped dither offsets fails here; the walk's step rule is
    pinned sample by sample in the next test."""
    sc = preset("Integer Power", 64, 48)
    sc.first_step_random = 0
    p = O.params(sc)
    job = H.post_job(deao_quality=quality, deao_dither=dither, deao_first_step_random=rnd)
    r = O.render(p, job, threads=0)
    vg = np.array([[p.vgrads[i][k] for k in range(3)] for i in range(3)])
    axes = vg / np.linalg.norm(vg, axis=1, keepdims=True)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    rng = np.random.default_rng(quality + 10 * dither)
    gaps = []
    for i in rng.choice(surf, size=min(30, surf.size), replace=False):
        x, y = int(i % p.width), int(i // p.width)
        rec = r["gbuf"][i]
        if dither:
            base = _ao_set(quality, (y % (dither + 1)) * 0.5 / dither, (x % (dither + 1)) * 0.5 / dither)
        else:
            base = _ao_set(quality)
        st = _row_state(p, r["gbuf"], x, y, len(base)) if rnd else None
        n, amb, pr = O.deao_probe(p, job, x, y, rec, rnd=st)
        assert n == len(base) and amb == (rec[4] & 0xFFFF)
        if rnd:
            assert np.all((pr["first"] >= 0.5) & (pr["first"] <= 2.0)) and np.ptp(pr["first"]) > 0
        else:
            assert np.all(pr["first"] == 1.0)
        c, v = O.pixel_ray(p, x, y)
        assert np.linalg.norm(pr["point"] - (c + r["depth"][i] * v)) < 0.05 * p.step_width
        nq = np.array([np.int16(rec[0] & 0xFFFF), np.int16(rec[0] >> 16), np.int16(rec[1] & 0xFFFF)], float)
        nv = nq / np.linalg.norm(nq)
        rot = _rodrigues(np.array([0.0, 0.0, -1.0]), nv)
        want = (base @ rot.T) @ axes  # view frame, then world
        got = pr["dirs"] / np.linalg.norm(pr["dirs"], axis=1, keepdims=True)
        assert np.abs(got - want).max() < 1e-5, (x, y)
        # one step width long (the view's axes normalised to it, src/deao.rs:286)
        assert np.allclose(np.linalg.norm(pr["dirs"], axis=1), p.step_width, rtol=1e-6)
        # the openness of each ray against a fine walk of the DE along it
        for k in range(n):
            t0 = pr["step_ao"] * pr["dsm"] * pr["first"][k]
            t = np.concatenate([np.arange(t0, pr["max_dist"] + 1e-9, 0.05), [t0, pr["max_dist"]]])
            de, _ = O.de_many(p, pr["point"][None, :] + np.outer(t, got[k]) * p.step_width)
            walk = min(1.0, float(np.min((de - pr["ms_ao"] + t * pr["md_d10"]) / t)))
            walk = max(walk, 0.0)
            s = pr["min_ra"][k] / pr["de_mul"]
            assert s >= walk - 0.02, (x, y, k, s, walk)
            gaps.append(s - walk)
        if np.all(pr["min_ra"] / pr["de_mul"] >= 1.0 - 1e-6):
            assert amb == 0
    gaps = np.abs(np.array(gaps))
    assert np.median(gaps) <= 0.025 and (gaps < 0.02).mean() >= 0.5, (np.median(gaps), (gaps < 0.02).mean())
    assert gaps.max() <= 0.3, gaps.max()


@pytest.mark.parametrize("quality,rnd", [(1, False), (2, True), (3, False)])
def test_deao_walk_steps_by_the_larger_of_the_de_and_its_growth(quality, rnd):
    """The walk along each DE AO ray (src/deao.rs:357-382), sample by
    sample, against the DE evaluated apart at each sample's point (de_many,
    a fresh marcher on the point surface + t * ray step):
      - the first sample at step_ao * dsm (times the ray's random factor);
      - the walk's DE at each sample equals the DE there;
      - each next sample lies the larger of that DE and dsm times the
        distance walked further on (src/deao.rs:379), held at max_dist;
      - the walk ends at the first sample whose running openness
        min((DE - stop + t md) / t) falls under 0.02, or at max_dist, and
        the ray's openness is that running minimum;
      - the samples at most 2 + log(max_dist / t0) / log(1 + dsm): the
        walk grows at least geometrically.
    The smaller of the two instead of the larger, a step of the DE alone or
    of the growth alone, a doubled step, an assignment in place of the
    increment, or a walk not held at max_dist fails here; the openness
    the walk reaches is held to a fine walk in
    test_deao_against_brute_force."""
    sc = preset("Integer Power", 64, 48)
    sc.first_step_random = 0
    p = O.params(sc)
    job = H.post_job(deao_quality=quality, deao_dither=0, deao_first_step_random=rnd)
    r = O.render(p, job, threads=0)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    rng = np.random.default_rng(100 + quality)
    base = _ao_set(quality)
    checked = long_walks = geometric = 0
    for i in rng.choice(surf, size=min(12, surf.size), replace=False):
        x, y = int(i % p.width), int(i // p.width)
        rec = r["gbuf"][i]
        st = _row_state(p, r["gbuf"], x, y, len(base)) if rnd else None
        n, _, pr = O.deao_probe(p, job, x, y, rec, rnd=st)
        for k in range(n):
            t, de, ms_walk, opn = O.deao_walk(p, job, x, y, rec, k, rnd=st)
            assert t.size >= 1
            t0 = pr["step_ao"] * pr["dsm"] * pr["first"][k]
            assert t[0] == pytest.approx(t0, rel=1e-12)
            pts = pr["point"][None, :] + t[:, None] * pr["dirs"][k][None, :]
            de_apart, _ = O.de_many(p, pts)
            floor = 0.25 * max(ms_walk, float(p.de_stop))
            free = de_apart > floor
            assert np.allclose(de[free], de_apart[free], rtol=1e-12, atol=0.0), (x, y, k)
            assert np.all(de[~free] <= floor * (1 + 1e-6))
            dsm, dmax = pr["dsm"], pr["max_dist"]
            for j in range(t.size - 1):
                want = t[j] + max(de[j], t[j] * dsm)
                if want > dmax:
                    want = dmax
                assert t[j + 1] == pytest.approx(want, rel=1e-12), (x, y, k, j)
                geometric += de[j] < t[j] * dsm
            run = np.minimum.accumulate(np.minimum(1.0, (de - pr["ms_ao"] + t * pr["md_d10"]) / t))
            assert np.all(run[:-1] >= 0.02) and np.all(t[:-1] < dmax), (x, y, k)
            assert run[-1] < 0.02 or t[-1] == dmax
            assert max(run[-1], 0.0) == pytest.approx(opn, rel=1e-6, abs=1e-7)
            assert t.size <= 2 + np.log(dmax / t0) / np.log1p(dsm)
            checked += 1
            long_walks += t.size >= 4
    # both branches of the step rule and long walks were exercised
    assert checked >= 30 and long_walks >= checked // 4 and geometric > 0, (checked, long_walks, geometric)


def test_deao_generator_is_the_c_runtime_lcg():
    """get_rand's generator (src/deao.rs:262-265) is the linear congruential
    generator of the Microsoft C runtime's rand(): state' = 214013 state +
    2531011 mod 2^32 (0x343FD, 0x269EC3), whose outputs, bits 16..30 of the
    state, from srand(1) are the published 41, 18467, 6334, 26500, ...; the
    AO's number is bits 8..30 of the state over 2^23 - 1, and a row's seed
    is 0x24563487 + (y + 1) 0x324594A1 wrapped to 32 bits
    (src/deao.rs:254-262)."""
    states, vals = O.deao_rand(1, 10)
    u = states.astype(np.int64) & 0xFFFFFFFF
    assert list((u >> 16) & 0x7FFF) == [41, 18467, 6334, 26500, 19169, 15724, 11478, 29358, 26962, 24464]
    assert np.array_equal(vals * 0x7FFFFF, ((u >> 8) & 0x7FFFFF).astype(float))
    assert vals.min() >= 0.0 and vals.max() <= 1.0
    for y in (0, 1, 1079, 4000):
        assert O.deao_row_seed(y) & 0xFFFFFFFF == (0x24563487 + (y + 1) * 0x324594A1) & 0xFFFFFFFF


@pytest.mark.parametrize("dither", [0, 1])
def test_deao_random_first_steps_draw_along_the_row(dither):
    """deao_row draws the random first steps from one generator per row,
    each surface pixel left to right taking one number per ray
    (src/deao.rs:262-270, 348-351): every surface pixel of three rows gets
    the amb word of the probe started from the row's seed advanced over the
    pixels before it, and its rays' factors are 1.5 u + 0.5 of the draws.
    A generator seeded per pixel, or a draw skipped on a row, fails here."""
    sc = preset("Integer Power", 64, 48)
    sc.first_step_random = 0
    p = O.params(sc)
    job = H.post_job(deao_quality=1, deao_dither=dither, deao_first_step_random=True)
    r = O.render(p, job, threads=0)
    plain = O.render(p, H.post_job(deao_quality=1, deao_dither=dither), threads=0)
    g = r["gbuf"]
    obj = _object(g)
    amb = g[:, 4] & 0xFFFF
    # the random steps change the picture
    assert (amb[obj] != (plain["gbuf"][:, 4] & 0xFFFF)[obj]).mean() > 0.2
    checked = 0
    for y in (10, 24, 37):
        st = O.deao_row_seed(y)
        for x in range(p.width):
            i = y * p.width + x
            if not obj[i]:
                continue
            n, a, pr = O.deao_probe(p, job, x, y, g[i], rnd=st)
            assert a == amb[i], (x, y)
            _, u = O.deao_rand(st, n)
            assert np.array_equal(pr["first"], u * 1.5 + 0.5)
            st = pr["rnd"]
            checked += 1
    assert checked > 20


def test_deao_random_first_step_past_the_ray_end():
    """With rays shorter than their first step (deao_max_len 0.02), a ray
    samples its random first step where it falls, not held to the ray's
    length, and then the ray's end (src/deao.rs:348-356): its openness is
    min(1, o(t0), o(max_dist)) with o(t) = (DE(t) - stop + t md_d10) / t,
    the second sample skipped when the first is under 0.02."""
    sc = preset("Integer Power", 64, 48)
    sc.first_step_random = 0
    p = O.params(sc)
    job = H.post_job(deao_quality=1, deao_first_step_random=True)
    job.deao_max_len = 0.02
    r = O.render(p, job, threads=0)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    past = 0
    for i in np.random.default_rng(5).choice(surf, size=12, replace=False):
        x, y = int(i % p.width), int(i // p.width)
        st = _row_state(p, r["gbuf"], x, y, 7)
        n, amb, pr = O.deao_probe(p, job, x, y, r["gbuf"][i], rnd=st)
        assert amb == (r["gbuf"][i][4] & 0xFFFF)
        unit = pr["dirs"] / np.linalg.norm(pr["dirs"], axis=1, keepdims=True)
        for k in range(n):
            t0 = pr["step_ao"] * pr["dsm"] * pr["first"][k]
            if t0 <= pr["max_dist"]:
                continue
            past += 1
            t = np.array([t0, pr["max_dist"]])
            de, _ = O.de_many(p, pr["point"][None, :] + np.outer(t, unit[k]) * p.step_width)
            o = (de - pr["ms_ao"] + t * pr["md_d10"]) / t
            s = min(1.0, o[0])
            if s >= 0.02:
                s = min(s, o[1])
            assert abs(pr["min_ra"][k] / pr["de_mul"] - max(s, 0.0)) < 1e-6, (x, y, k)
    assert past > 20
