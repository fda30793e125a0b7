"""The oracle's secondary-ray march (orc_march_rays: SrTracer::march_ray,
src/reflect.rs:1464-1674) pinned on the CPU.

- A ray that starts on a pixel's start plane along the pixel's own view
  ray (zzplus = 1, depth 0) is the primary march with another first step
  and bisection: it meets the surface the picture shows, at its depth to
  the bisection's tolerance, with the same normal.
- Rays that start inside the set, that leave the set's bounding sphere, or
  that are cut short by max_len end in the open air, at the far end or at
  their limit.
- With depth 0 at the start the square-root depth code (the painter's far
  fog) is the plain one.
- The result does not depend on the thread count.
- The hits' shadow and occlusion pass (calc_hs and deao_1pos,
  src/reflect.rs:805-1090), on view rays that reproduce the primary hits:
  the hard-shadow bit of a light equals the primary shadow pass's for the
  same light at >= 97 % of the pixels, the AO follows the primary DE AO
  (correlation >= 0.9; the two differ by the primary's neighbour
  correction), and the soft shadow is darker where the hard one is set.
"""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O


def pixel_rays(p):
    w, h = p.width, p.height
    st = np.zeros((w * h, 3))
    vf = np.zeros((w * h, 3))
    for y in range(h):
        for x in range(w):
            st[y * w + x], vf[y * w + x] = O.pixel_ray(p, x, y)
    return st, vf


def view_tasks(p, st, vf):
    n = p.width * p.height
    t = np.zeros(n, H.RAY_TASK_DTYPE)
    t["origin"] = st
    t["dir"] = vf.astype(np.float32)
    t["pixel"] = np.arange(n, dtype=np.uint32)
    return t


@pytest.mark.parametrize("preset,scale", [("Integer Power", 2.0), ("Amazing Box", 4.0)])
def test_a_view_ray_is_the_primary_march(preset, scale):
    sc = H.scene_preset(preset, 96, 72)
    p = H.params_from_scene(sc)
    r = O.render(p, None, threads=0)
    st, vf = pixel_rays(p)
    t = view_tasks(p, st, vf)
    # the task's direction is f32: march the picture's rays with the same
    # rounding so that both follow one line
    hh = O.march_rays(p, t)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    hit = (hh["flags"] & H.RAY_HIT) != 0
    assert (surf == hit).mean() >= 0.99, (surf.mean(), hit.mean())
    both = surf & hit
    assert both.sum() > 500
    # depth along the view ray: zz (= zz2 here) against the march's mzz,
    # in steps.  The two marches sample the ray at other places (another
    # first step), so their bisections end at other points of the same
    # crossing: thousandths of a step, more only at silhouettes
    dz = np.abs(hh["zz"][both] - r["depth"][both])
    assert np.median(dz) < 0.01 and (dz < 0.1).mean() >= 0.98, np.percentile(dz, [50, 90, 99, 100])
    assert np.allclose(hh["zz"][both], hh["zz2"][both], rtol=1e-9, atol=1e-9)
    ng = hh["gbuf"][both].astype(np.int64)
    no = r["gbuf"][both].astype(np.int64)
    def nrm(g):
        v = np.stack([(g[:, 0] & 0xFFFF), g[:, 0] >> 16, g[:, 1] & 0xFFFF], axis=1)
        v = np.where(v >= 32768, v - 65536, v).astype(np.float64)
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    # the same normal, to the depth difference: a fractal's normal turns
    # within a fraction of a step, so a few per cent differ by more than 2
    # degrees (a wrong axis order or sign would turn most of them)
    cos = np.sum(nrm(ng) * nrm(no), axis=1)
    assert np.median(cos) > np.cos(np.radians(0.1)), np.percentile(cos, [1, 50])
    assert (cos > np.cos(np.radians(2.0))).mean() >= 0.95, np.percentile(cos, [1, 50])


def test_open_air_rays():
    sc = H.scene_preset("Integer Power", 64, 48)
    p = H.params_from_scene(sc)
    rng = np.random.default_rng(3)
    n = 600
    t = np.zeros(n, H.RAY_TASK_DTYPE)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # outside the bounding sphere (radius 2), heading outwards
    t["origin"] = d * 3.0
    t["dir"] = (d * p.step_width).astype(np.float32)
    t["pixel"] = rng.integers(0, 64 * 48, n)
    t["zz"] = 100.0
    hh = O.march_rays(p, t)
    assert not (hh["flags"] & H.RAY_HIT).any()
    assert np.all(hh["zz2"] > p.zend) and np.all(hh["gbuf"][:, 2] == 32768 << 16)
    assert np.all(hh["gbuf"][:, 1] >> 16 == np.round(np.clip(hh["step_count"], 0, 1023)).astype(np.uint32))
    # inside: no march at all
    t["inside"] = 1
    hi = O.march_rays(p, t)
    assert not (hi["flags"] & H.RAY_HIT).any() and np.all(hi["zz2"] == p.zend) and np.all(hi["step_count"] == 0)


def test_max_len_cuts_the_ray():
    sc = H.scene_preset("Integer Power", 64, 48)
    p = H.params_from_scene(sc)
    st, vf = pixel_rays(p)
    t = view_tasks(p, st, vf)
    full = O.march_rays(p, t)
    hit = (full["flags"] & H.RAY_HIT) != 0
    lim = np.float32(np.median(full["zz2_paint"][hit]))
    t["max_len"] = lim
    cut = O.march_rays(p, t)
    ch = (cut["flags"] & H.RAY_HIT) != 0
    # what stopped short of the limit is the same hit; the rest is cut
    short = hit & (full["zz2_paint"] <= lim)
    assert np.array_equal(ch, short)
    assert np.all(cut["zz2"][~ch & hit] > lim)
    assert np.array_equal(cut["gbuf"][ch], full["gbuf"][ch])


def test_square_root_depth_is_the_plain_one_from_depth_zero():
    sc = H.scene_preset("Amazing Box", 64, 48)
    p = H.params_from_scene(sc)
    st, vf = pixel_rays(p)
    t = view_tasks(p, st, vf)
    a = O.march_rays(p, t, depth_sqrt=False)
    b = O.march_rays(p, t, depth_sqrt=True)
    assert np.array_equal(a["gbuf"], b["gbuf"])
    t["zz"] = 50.0
    c = O.march_rays(p, t, depth_sqrt=True)
    d = O.march_rays(p, t, depth_sqrt=False)
    hit = (c["flags"] & H.RAY_HIT) != 0
    # a larger depth value gives a smaller depth code (nearer is larger)
    assert np.all((c["gbuf"][hit, 2] >> 8) <= (d["gbuf"][hit, 2] >> 8))


def test_rays_do_not_depend_on_the_thread_count():
    sc = H.scene_preset("Integer Power", 64, 48)
    p = H.params_from_scene(sc)
    r = O.render(p, None, threads=0)
    st, vf = pixel_rays(p)
    t = H.reflection_tasks_from(p, st, vf, {"depth64": r["depth"], "normals": r["normals"]})
    assert len(t) > 300
    a = O.march_rays(p, t, threads=1)
    b = O.march_rays(p, t, threads=7)
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("preset", ["Integer Power", "Amazing Box"])
def test_the_hit_pass_follows_the_primary_shadows_and_occlusion(preset):
    sc = H.scene_preset(preset, 96, 72)
    p = H.params_from_scene(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], deao_quality=1, params=p)
    r = O.render(p, job, threads=0)
    st, vf = pixel_rays(p)
    t = view_tasks(p, st, vf)
    t["ao_q"] = 2  # 7 rays, as the primary quality 1
    o = H.RayOpts()
    o.n_lights, o.light_idx[0], o.hs_mul = 1, 0, job.shadow_max_len_mul
    o.light_vec[0][:] = list(job.light_vec[0])
    o.ao, o.ao_max_len = 1, job.deao_max_len
    hh = O.march_rays(p, t, opts=o)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    both = surf & ((hh["flags"] & H.RAY_HIT) != 0)
    a = (hh["gbuf"][both, 1] >> 16) & 0x400
    b = (r["gbuf"][both, 1] >> 16) & 0x400
    assert (a == b).mean() >= 0.97 and 0.05 < (a != 0).mean() < 0.95
    ao_r, ao_p = hh["gbuf"][both, 4] & 0xFFFF, r["gbuf"][both, 4] & 0xFFFF
    assert np.corrcoef(ao_r, ao_p)[0, 1] >= 0.9
    # the step counts and depths of the march are kept
    plain = O.march_rays(p, t)
    keep = np.array([0xFFFFFFFF, 0x0000FFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0], np.uint32)
    assert np.array_equal(hh["gbuf"] & keep, plain["gbuf"] & keep)
    # one soft shadow of the light: its value (0..63) is lower where the hard bit is set
    o.soft_radius = 2.0
    hs = O.march_rays(p, t, opts=o)
    soft = (hs["gbuf"][both, 1] >> 26).astype(np.float64)
    assert soft[a != 0].mean() + 10 < soft[a == 0].mean()
