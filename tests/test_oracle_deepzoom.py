"""The hybrid chain's deep zoom (Mandelbulb p8 x1, Amazing Box scale 2
x2, 45 iterations, the reference's numeric DE), on the oracle (CPU).

  - The chain's orbits stay bounded only in a core around the origin
    (radius ~0.006-0.008): there the box steps scale by 8 (inside MinR, the
    sphere fold's 4 times the scale 2) and the p8 step shrinks.  Its DE
    shell, where the normal-zoom pictures stop (a pixel, 0.067 units, out),
    holds no bounded orbit: along the coarse picture's surface rays no
    point's orbit stays bounded, and closer views find the DE no smaller.
    So config 2's camera sits 1e-9 outside the core's surface, found by a
    seeded coarse march into it and bisected (mb3d_hip.scene_deepzoom).
  - At 1e-9 the numeric DE's offset, 0.004 of a step width (~3e-17
    units), is under the f64 resolution of coordinates near 1 (2.2e-16)
    but ~37 ulps of the core's coordinates (~0.007): the DE's finite
    differences resolve there, and the oracle's picture is the surface."""
import math

import numpy as np

import mb3d_hip as H
import oracle as O


def test_the_chain_has_no_bounded_orbits_along_its_surface_rays():
    sc = H.scene_hybrid(64, 48, seed=1)
    sc.bin_search_steps = 30
    p = O.params(sc)
    r = O.render(p, threads=0)
    hits = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    assert hits.size > 100
    top = 0
    for i in hits[::23][:12]:
        c, v = O.pixel_ray(p, i % 64, i // 64)
        u = v / np.linalg.norm(v)
        ts = r["depth"][i] * np.linalg.norm(v) + np.linspace(-0.05, 0.3, 4001)
        _, its = O.de_many(p, c[None, :] + ts[:, None] * u[None, :])
        top = max(top, int(its.max()))
    assert top < p.max_it // 2, top


def test_closer_views_find_the_chain_no_nearer():
    sc = H.scene_hybrid(64, 48, seed=1)
    p = O.params(sc)
    r = O.render(p, threads=0)
    hits = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    i = int(hits[len(hits) // 2])
    c, v = O.pixel_ray(p, i % 64, i // 64)
    surface = c + r["depth"][i] * v
    u = v / np.linalg.norm(v)
    for dist in (0.1, 0.03):
        ts = np.linspace(0.0, 2.0 * dist, 2001)
        pts = (surface - dist * u)[None, :] + ts[:, None] * u[None, :]
        de = np.array([O.de_at(p, q)[0] for q in pts[::20]])  # world units
        assert de.min() > 0.01, (dist, de.min())


def test_the_numeric_offset_is_below_f64_resolution_at_1e_9():
    sc = H.scene_hybrid(1920, 1080, seed=1)
    sw = 1e-9 * math.radians(0.5) / 1080  # the step width of the 1e-9 view (FOV 0.5)
    sc.zoom = 2.1345 / (1920 * sw)
    p = O.params(sc)
    assert p.de_offset < 0.5 * np.spacing(1.0), p.de_offset


def test_the_chain_is_bounded_only_in_a_core_around_the_origin():
    p = O.params(H.scene_hybrid(64, 48, seed=1))
    rng = np.random.default_rng(7)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    _, its_in = O.de_many(p, 0.004 * d * rng.random((400, 1)))
    assert (its_in >= p.max_it).all()
    _, its_out = O.de_many(p, 0.1 * d + 0.9 * d * rng.random((400, 1)))
    assert (its_out < p.max_it).all()


def test_the_deep_zoom_camera_faces_the_core_surface():
    sc = H.scene_deepzoom(96, 54, 1, O.de_many)
    p = O.params(sc)
    surf, u = H.chain_surface_point(1, O.de_many)
    assert 0.004 < np.linalg.norm(surf) < 0.05
    _, its = O.de_many(p, np.stack([surf, surf - 1e-15 * u]))
    assert its[0] < p.max_it <= its[1]  # escaping side, then inside
    assert p.de_offset > 8 * np.spacing(np.abs(surf).max())  # the DE's differences resolve
    r = O.render(p, threads=0)
    d = r["depth"]
    hit = np.isfinite(d) & (d > 0)
    assert hit.mean() > 0.9
    # the surface 1e-9 ahead: depths of ~1e-9 scene units
    t = np.median(d[hit]) * p.step_width
    assert 0.5e-9 < t < 2e-9, t
