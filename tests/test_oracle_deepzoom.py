"""Why the hybrid chain's deep zoom (Mandelbulb p8 x1, Amazing Box scale 2
x2, 45 iterations, the reference's numeric DE) has no surface at 1e-9: the
oracle's own evidence (CPU).

  - Along the coarse picture's surface rays, no point's orbit stays bounded
    for the 45 iterations: the set the iteration limit draws is dust that
    fine sampling does not meet, so no march stops on the limit.
  - The coarse picture's surface is where the DE falls under its stop (a
    pixel, 0.067 scene units).  Closer views along the same ray find the DE
    no smaller: the estimate keeps the set tens of thousandths of a unit
    away, so a view 1e-9 from that point is background, for the reference's
    march as for this backend's (tests/test_gpu_deepzoom.py holds the two
    pictures equal).
  - And at 1e-9 the numeric DE's offset, 0.004 of a step width (about 3e-17
    units), is under the f64 resolution of the coordinates (2.2e-16 at 1):
    its finite differences vanish."""
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
