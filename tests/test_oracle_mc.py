"""The oracle's Monte Carlo marches (orc_mc_batch: McMarch::run,
src/mc.rs:488-977) pinned on the CPU.

- The random numbers are GetRand's (src/mc.rs:470-476): the linear
  congruential generator of the C runtime's rand, whose outputs from seed 1
  (state >> 16 & 0x7FFF) are the well-known 41, 18467, 6334, ...; the
  returned seed counts the draws, and the first step of a ray is its full
  step times the draw.
- A ray along a pixel's view ray from its start plane (depth 0, zzplus 1)
  meets the surface the picture shows at the picture's depth, with the
  picture's normal (to the one-sided stencil's error), gradient and fog
  step count.
- The normal is the one-sided difference of the DE on the world axes, at
  min(DE stop, 1) (1 + |zz| DE stop factor) 0.15 steps, on the side bit 7
  of the seed picks: recomputed from the oracle's DE at points.
- The secant search and the bisection end at the same crossing.
- A primary ray whose start is inside the set returns the picture's
  inside-start colour word (src/calc.rs:2277-2293) bit for bit.
- A shadow ray towards a light is blocked where the picture's hard shadow
  pass (calc_hs) sets the light's bit.
- With the inside flag the DE is turned around (invert_inside): a ray from
  inside the set stops where it leaves it; the way through a transparent
  object ends on the set's last crossing of the segment.
- The result does not depend on the thread count.
"""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

C_RAND_FROM_1 = [41, 18467, 6334, 26500, 19169, 15724, 11478, 29358, 26962, 24464]


def view_tasks(p, kind=H.MC_RAY):
    """Every pixel's view ray from its start plane (one step long)."""
    w, h = p.width, p.height
    n = w * h
    t = np.zeros(n, H.MC_TASK_DTYPE)
    for y in range(h):
        for x in range(w):
            t["pos"][y * w + x], t["dir"][y * w + x] = O.pixel_ray(p, x, y)
    t["kind"] = kind
    t["zzplus"] = 1.0
    t["len"] = p.zend
    t["seed"] = (np.arange(n, dtype=np.int64) * 7919 + 13).astype(np.int32)
    return t


def lcg(seed):
    """One draw of the C runtime's rand on int32 seeds."""
    s = np.asarray(seed).astype(np.int64) & 0xFFFFFFFF
    return ((s * 214013 + 2531011) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def world_normals(p, n_view):
    """The picture's normals (on the camera's axes) on the world axes."""
    vg = np.array([[p.vgrads[i][k] for k in range(3)] for i in range(3)])
    w = n_view.reshape(-1, 3) @ (vg / np.linalg.norm(vg, axis=1, keepdims=True))
    return w / np.maximum(np.linalg.norm(w, axis=1, keepdims=True), 1e-300)


def test_the_random_numbers_are_the_c_runtime_rand():
    p = H.params_from_scene(H.scene_preset("Integer Power", 32, 24))
    t = view_tasks(p)[:1].copy()
    # a ray cut at length 0 ends after its first step: one draw
    t["len"] = 0.0
    s, seq = 1, []
    for _ in range(10):
        t["seed"] = s
        r = O.mc_batch(p, t)
        assert r["kind"][0] == 0
        s = int(r["seed"][0])
        seq.append(((s & 0xFFFFFFFF) >> 16) & 0x7FFF)
    assert seq == C_RAND_FROM_1
    # the first step is the full step times the draw (state & 0x7FFFFFFF) / 0x7FFFFFFF
    t2 = np.repeat(t, 2)
    t2["seed"] = [1, 987654321]
    r = O.mc_batch(p, t2)
    u = (r["seed"].astype(np.int64) & 0x7FFFFFFF) / 0x7FFFFFFF
    assert np.all(r["zz2"] > 0)
    assert r["zz2"][0] / r["zz2"][1] == pytest.approx(u[0] / u[1], rel=1e-12)


@pytest.mark.parametrize("preset", ["Integer Power", "Amazing Box"])
def test_a_view_ray_meets_the_picture(preset):
    p = H.params_from_scene(H.scene_preset(preset, 96, 72))
    r = O.render(p, None, threads=0)
    t = view_tasks(p)
    m = O.mc_batch(p, t)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    hit = m["kind"] == 1
    assert (surf == hit).mean() >= 0.995, (surf.mean(), hit.mean())
    both = surf & hit
    assert both.sum() > 1000
    dz = np.abs(m["zz"][both] - r["depth"][both])
    assert np.median(dz) < 0.01 and (dz < 0.1).mean() >= 0.98, np.percentile(dz, [50, 90, 99])
    # along the view ray the depth is the length: the bisection starts half the last step back from where
    # the march stopped (zz2)
    assert (m["zz"][both] < m["zz2"][both] + 0.05).mean() >= 0.99
    # the normal: unit, facing the ray, within the one-sided stencil's error of the picture's centred one
    nm = m["normals"][both]
    assert np.allclose(np.linalg.norm(nm, axis=1), 1.0, atol=1e-12)
    assert np.all(np.sum(nm * t["dir"][both], axis=1) < 0.0)
    deg = np.degrees(np.arccos(np.clip(np.sum(world_normals(p, r["normals"])[both] * nm, axis=1), -1, 1)))
    assert np.median(deg) < 4.0 and (deg < 15.0).mean() >= 0.9, np.percentile(deg, [50, 90, 99])
    # the gradient of the same smoothed iteration count and DE stop (the stencil's last DE gives the
    # orbit trap, another point than the picture's)
    dg = np.abs((r["gbuf"][both, 3] & 0xFFFF).astype(np.int64) - m["si_grad"][both])
    assert np.median(dg) <= 2 and (dg <= 10).mean() >= 0.9, np.percentile(dg, [50, 90, 99])
    assert np.all((m["flags"][both] & H.MC_DE_LIMITED) != 0) or preset != "Amazing Box"
    # the fog's step count: the picture's less its random first step, to a step or two
    df = m["dyn_fog"][both] - (r["gbuf"][both, 1] >> 16)
    assert (np.abs(df) <= 2.5).mean() >= 0.98, np.percentile(df, [1, 50, 99])
    # open air: past the far end, the position the last step left
    air = m["kind"] == 0
    assert np.all(m["zz2"][air] > p.zend)
    assert np.all(m["seed"][air] == lcg(t["seed"][air]))


def test_the_normal_is_the_one_sided_world_stencil():
    p = H.params_from_scene(H.scene_preset("Integer Power", 64, 48))
    t = view_tasks(p)
    m = O.mc_batch(p, t)
    s1 = lcg(t["seed"])
    once = (m["kind"] == 1) & (m["seed"] == s1)  # the first step's draw only: no second try
    assert once.sum() > 300 and (m["kind"] == 1).sum() - once.sum() > 0
    base = float(min(np.float32(p.de_stop), np.float32(1.0)))
    for i in np.flatnonzero(once)[::4][:120]:
        sca = float(np.float32(-0.15 if int(s1[i]) & 128 else 0.15))
        off = base * (1.0 + abs(m["zz"][i]) * float(p.de_stop_factor)) * sca * p.step_width
        c = m["pos"][i].copy()
        d0 = O.de_at(p, c)[0]
        n = np.zeros(3)
        for k in range(3):
            e = c.copy()
            e[k] = c[k] + off
            n[k] = (O.de_at(p, e)[0] - d0) * sca
        n /= np.linalg.norm(n)
        assert np.dot(n, m["normals"][i]) > 1.0 - 1e-12, (i, n, m["normals"][i])


def test_secant_search_and_bisection_end_at_one_crossing():
    p = H.params_from_scene(H.scene_preset("Amazing Box", 64, 48))
    t = view_tasks(p)
    a = O.mc_batch(p, t, secant=False)
    b = O.mc_batch(p, t, secant=True)
    assert np.array_equal(a["kind"], b["kind"]) and np.array_equal(a["zz2"], b["zz2"])
    hit = a["kind"] == 1
    dz = np.abs(a["zz"][hit] - b["zz"][hit])
    assert np.median(dz) < 0.005 and (dz < 0.2).mean() >= 0.98, np.percentile(dz, [50, 98, 100])
    assert not np.array_equal(a["zz"][hit], b["zz"][hit])


def test_a_primary_ray_that_starts_inside():
    sc = H.scene_preset("Integer Power", 64, 48)
    sc.z_start = 0.0  # the start plane through the middle of the set
    p = H.params_from_scene(sc)
    r = O.render(p, None, threads=0)
    t = view_tasks(p)
    t["flags"] = H.MC_PROBE
    m = O.mc_batch(p, t)
    ins = r["depth"] == 0
    assert 0.05 < ins.mean() < 0.95, ins.mean()
    assert np.array_equal(m["kind"] == 2, ins)
    assert np.array_equal(m["si_grad"][ins], r["gbuf"][ins, 3] & 0xFFFF)
    assert np.array_equal(m["otrap"][ins], r["gbuf"][ins, 3] >> 16)
    assert np.all(m["si_grad"][ins] >= 32768) and np.all(m["seed"][ins] == t["seed"][ins])
    # without the probe the same rays march
    t["flags"] = 0
    assert not np.any(O.mc_batch(p, t)["kind"] == 2)


def test_shadow_rays_follow_the_hard_shadow_pass():
    p = H.params_from_scene(H.scene_preset("Integer Power", 96, 72))
    light = (0.6, -0.5, -0.6)
    job = H.post_job(lights=[light], params=p)
    r = O.render(p, job, threads=0)
    t = view_tasks(p)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    idx = np.flatnonzero(surf)
    s = t[idx].copy()
    d = r["depth"][idx]
    s["pos"] = t["pos"][idx] + t["dir"][idx] * d[:, None]
    v = -np.array(job.light_vec[0][:])
    vf = t["dir"][idx]
    s["dir"] = v
    s["kind"] = H.MC_SHADOW
    s["zz"] = s["zz2"] = d
    s["zzplus"] = vf @ v / np.sqrt((vf * vf).sum(axis=1) * (v @ v))
    s["len"] = 300.0
    m = O.mc_batch(p, s)
    assert np.all(m["kind"] == 0) and np.all(m["pos"] == 0.0)
    blocked = m["len"] > 0.01
    hard = ((r["gbuf"][idx, 1] >> 16) & 0x400) != 0
    print("blocked", blocked.mean(), "hard", hard.mean(), "agree", (blocked == hard).mean())
    assert 0.05 < hard.mean() < 0.95
    assert (blocked == hard).mean() >= 0.94
    # with `trans` a blocked ray returns the surface it met, facing it
    s["flags"] = H.MC_TRANS
    mt = O.mc_batch(p, s)
    assert np.array_equal(mt["kind"] == 1, blocked) and np.array_equal(mt["len"], m["len"])
    b = mt[blocked]
    assert np.all(np.sum(b["normals"] * v, axis=1) < 0.0) and np.all(b["ms_de_stop"] > 0)
    de = np.array([O.de_at(p, q)[0] for q in b["pos"][:200]]) / p.step_width
    assert np.median(np.abs(de - b["ms_de_stop"][:200])) < 0.05 * np.median(b["ms_de_stop"])


def test_inside_rays_stop_where_they_leave_the_set():
    p = H.params_from_scene(H.scene_preset("Integer Power", 48, 36))
    r = O.render(p, None, threads=0)
    t = view_tasks(p)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    s = t[surf].copy()
    s["pos"] = t["pos"][surf] + t["dir"][surf] * (r["depth"][surf] + 3.0)[:, None]
    s["zz"] = s["zz2"] = r["depth"][surf] + 3.0
    s["flags"] = H.MC_INSIDE
    s["len"] = s["zz"] + 400.0
    m = O.mc_batch(p, s)
    inside_start = np.array([O.de_at(p, q)[0] for q in s["pos"]]) / p.step_width < p.de_stop
    hit = m["kind"] == 1
    assert np.all(m["flags"] & H.MC_INSIDE)
    print("inside starts", inside_start.mean(), "hits", hit.mean())
    # the ray ends where the plain DE reaches the DE stop again: on the set's far side of the start
    k = np.flatnonzero(hit & inside_start)
    assert len(k) > 50
    de_end = np.array([O.de_at(p, q)[0] for q in m["pos"][k]]) / p.step_width
    ms = p.de_stop * (1.0 + np.abs(m["zz"][k]) * p.de_stop_factor)
    assert np.median(np.abs(de_end - ms) / ms) < 0.1, np.percentile(np.abs(de_end - ms) / ms, [50, 90])
    # the normal of the exit faces back along the ray (against the plain gradient)
    assert np.all(np.sum(m["normals"][k] * s["dir"][k], axis=1) < 0.0)


def _last_crossing(p, st, vf, a, b, ms):
    """The farthest length in [a, b] at which the plain DE is below ms, on a
    grid of a quarter step, refined by bisection (nan: none)."""
    ks = np.arange(a, b, 0.25)
    de = np.array([O.de_at(p, st + vf * k)[0] for k in ks]) / p.step_width
    inn = np.flatnonzero(de < ms)
    if len(inn) == 0 or inn[-1] + 1 >= len(ks):
        return np.nan
    lo, hi = ks[inn[-1]], ks[inn[-1] + 1]
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        if O.de_at(p, st + vf * mid)[0] / p.step_width < ms:
            lo = mid
        else:
            hi = mid
    return lo


def test_the_way_through_ends_on_the_last_crossing():
    """From just beyond the set's last crossing of the view ray back towards
    the start: the walk never stops in the open part of the segment, and it
    ends on the set after a first step of 1 to 9 and then steps of 1 + 3 DE
    (overshooting into the set by up to twice the DE)."""
    p = H.params_from_scene(H.scene_preset("Integer Power", 48, 36))
    r = O.render(p, None, threads=0)
    t = view_tasks(p, kind=H.MC_THROUGH)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))[::4][:100]
    s = t[surf].copy()
    ms = float(p.de_stop)
    last = np.array([_last_crossing(p, s["pos"][i], s["dir"][i], r["depth"][surf[i]] - 1.0,
                                    r["depth"][surf[i]] + 120.0, ms) for i in range(len(s))])
    ok = np.isfinite(last)
    assert ok.mean() > 0.9
    s, last = s[ok], last[ok]
    rng = np.random.default_rng(5)
    far = last + rng.uniform(0.5, 3.0, len(s))
    s["len"] = far
    s["flags"] = H.MC_INSIDE
    s["ms_de_stop"] = ms
    m = O.mc_batch(p, s)
    assert np.all(m["kind"] == 0) and np.all(m["seed"] == lcg(s["seed"]))
    d = m["len"] - last
    assert np.all(d <= 0.0), np.percentile(d, [50, 99, 100])
    # (a step of 1 + 3 DE can jump a thin part of the set: a few walk on)
    assert (d >= -25.0).mean() >= 0.9 and np.median(d) > -10.0, np.percentile(d, [1, 50, 99])
    # where it stopped the plain DE is within the DE stop: on the set
    de = np.array([O.de_at(p, s["pos"][i] + s["dir"][i] * m["len"][i])[0] for i in range(len(s))]) / p.step_width
    assert np.all(de[m["len"] > 0] < ms)
    # without the inside flag the DE is not turned around: the walk goes on through the set and stops
    # off it
    s["flags"] = 0
    o = O.mc_batch(p, s)
    de = np.array([O.de_at(p, s["pos"][i] + s["dir"][i] * o["len"][i])[0] for i in range(len(s))]) / p.step_width
    assert np.all(de[o["len"] > 0] > ms) and np.all(o["len"] <= far - 1.0)


def test_the_results_do_not_depend_on_the_thread_count():
    p = H.params_from_scene(H.scene_preset("Integer Power", 48, 36))
    t = view_tasks(p)
    t["kind"][1::3] = H.MC_SHADOW
    t["kind"][2::3] = H.MC_THROUGH
    t["flags"][::2] = H.MC_INSIDE
    t["len"][2::3] = 50.0
    a = O.mc_batch(p, t, threads=1)
    b = O.mc_batch(p, t, threads=7)
    assert a.tobytes() == b.tobytes()
    bad = t[:2].copy()
    bad["kind"][1] = 3
    with pytest.raises(RuntimeError):
        O.mc_batch(p, bad)
