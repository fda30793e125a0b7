"""The Monte Carlo renderer's marches on the device (mb3d_hip_mc_batch,
k_mc: McMarch::run, src/mc.rs:488-977, as gpu::mc_batch batches them,
src/gpu.rs:3013-3190) against the oracle on the same tasks.

Both sides march in f64 in the reference's operation order without FMA
contraction; they differ by the device's libm in the formulas that call
sin/cos/pow and in the gradient's log.  The bar per batch: result kind and
seed (the draws taken) equal for >= 99.9 % of the tasks; over the tasks
both end on a surface, the depth and the ray length within 1e-4 steps, the
normal within 0.5 degrees, gradient, orbit trap and fog step count equal,
each for >= 99.9 %; the lengths left (shadow, through) within 1e-4 steps."""
import ctypes as C

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
from test_oracle_mc import view_tasks

pytestmark = pytest.mark.gpu


def compare(a, b, what):
    n = len(a)
    m = {"tasks": n, "kind": (a["kind"] == b["kind"]).mean(), "seed": (a["seed"] == b["seed"]).mean(),
         "flags": (a["flags"] == b["flags"]).mean()}
    both = (a["kind"] == 1) & (b["kind"] == 1)
    k = int(both.sum())
    m["surfaces"] = k
    if k:
        cos = np.sum(a["normals"][both] * b["normals"][both], axis=1)
        m["zz"] = (np.abs(a["zz"][both] - b["zz"][both]) <= 1e-4).mean()
        m["zz2"] = (np.abs(a["zz2"][both] - b["zz2"][both]) <= 1e-4).mean()
        m["normals"] = (cos > np.cos(np.radians(0.5))).mean()
        m["grad"] = (a["si_grad"][both] == b["si_grad"][both]).mean()
        m["otrap"] = (a["otrap"][both] == b["otrap"][both]).mean()
        m["fog"] = (a["dyn_fog"][both] == b["dyn_fog"][both]).mean()
    same = a["kind"] == b["kind"]
    m["len"] = (np.abs(a["len"][same] - b["len"][same]) <= 1e-4).mean()
    print(what, m, flush=True)
    for key in ("kind", "seed", "flags", "len"):
        assert m[key] >= 0.999, (key, m)
    if k:
        for key in ("zz", "zz2", "normals", "grad", "otrap", "fog"):
            assert m[key] >= 0.999, (key, m)
    return m


def run_both(hip, p, t, secant=False):
    got = hip.mc_batch(p, t, secant)
    want = O.mc_batch(p, t, secant)
    return got, want


@pytest.mark.parametrize("preset", ["Integer Power", "Amazing Box", "Quaternion", "Bulbox"])
def test_view_rays_match_the_oracle(hip, preset):
    p = hip.params_from_scene(H.scene_preset(preset, 160, 120))
    t = view_tasks(p)
    got, want = run_both(hip, p, t)
    m = compare(got, want, preset)
    assert m["surfaces"] > 500
    c = H.counters()
    assert c["pixels"] == len(t) and c["de_evals_f64"] > 0 and c["hits"] == (got["kind"] == 1).sum()


def test_the_secant_search_and_the_analytic_bulb(hip):
    """The benchmark's power-8 bulb (analytic DE, the FP32 kinds' f64 path)
    refined by the secant search."""
    p = hip.params_from_scene(H.scene_bulb(160, 120, seed=2))
    t = view_tasks(p)
    got, want = run_both(hip, p, t, secant=True)
    compare(got, want, "bulb, secant")


def test_primary_rays_that_start_inside(hip):
    sc = H.scene_preset("Integer Power", 128, 96)
    sc.z_start = 0.0
    p = hip.params_from_scene(sc)
    t = view_tasks(p)
    t["flags"] = H.MC_PROBE
    t["flags"][::2] |= H.MC_INSIDE
    got, want = run_both(hip, p, t)
    compare(got, want, "probe")
    ins = want["kind"] == 2
    assert ins.mean() > 0.05
    assert np.array_equal(got["si_grad"][ins], want["si_grad"][ins])
    assert np.array_equal(got["otrap"][ins], want["otrap"][ins])


def test_shadow_rays_and_their_surfaces(hip):
    p = hip.params_from_scene(H.scene_preset("Integer Power", 160, 120))
    want_r = O.render(p, None, threads=0)
    t = view_tasks(p)
    surf = np.flatnonzero(np.isfinite(want_r["depth"]) & (want_r["depth"] > 0))
    s = t[surf].copy()
    d = want_r["depth"][surf]
    s["pos"] = t["pos"][surf] + t["dir"][surf] * d[:, None]
    v = np.array([0.6, -0.5, -0.6])
    v *= p.step_width / np.linalg.norm(v)
    s["dir"] = v
    s["kind"] = H.MC_SHADOW
    s["zz"] = s["zz2"] = d
    vf = t["dir"][surf]
    s["zzplus"] = vf @ v / np.sqrt((vf * vf).sum(axis=1) * (v @ v))
    s["len"] = 300.0
    s["flags"][::2] = H.MC_TRANS
    got, want = run_both(hip, p, s)
    m = compare(got, want, "shadow")
    assert m["surfaces"] > 200
    sh = want["kind"] == 1
    assert np.allclose(got["ms_de_stop"][sh], want["ms_de_stop"][sh], rtol=1e-6)


def test_inside_rays_and_the_way_through(hip):
    p = hip.params_from_scene(H.scene_preset("Amazing Box", 128, 96))
    r = O.render(p, None, threads=0)
    t = view_tasks(p)
    surf = np.flatnonzero(np.isfinite(r["depth"]) & (r["depth"] > 0))
    s = t[surf].copy()
    deep = r["depth"][surf] + 3.0
    s["pos"] = t["pos"][surf] + t["dir"][surf] * deep[:, None]
    s["zz"] = s["zz2"] = deep
    s["len"] = deep + 400.0
    s["flags"] = H.MC_INSIDE
    # every third task walks back through the set from 40 steps on
    s["kind"][::3] = H.MC_THROUGH
    s["len"][::3] = 40.0
    s["ms_de_stop"][::3] = p.de_stop
    got, want = run_both(hip, p, s)
    m = compare(got, want, "inside and through")
    assert m["surfaces"] > 200


def test_results_on_the_device(hip):
    """Tasks and results in device memory (the renderer keeps them there
    between its passes)."""
    import torch

    p = hip.params_from_scene(H.scene_preset("Integer Power", 96, 72))
    t = view_tasks(p)
    t["kind"][1::3] = H.MC_THROUGH
    t["len"][1::3] = 30.0
    host = hip.mc_batch(p, t)
    dt = torch.from_numpy(t.view(np.uint8).copy()).cuda()
    dr = torch.zeros(len(t) * H.MC_RES_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    H.check(H.load().mb3d_hip_mc_batch(C.byref(p), C.c_void_p(dt.data_ptr()), len(t), 0,
                                       C.c_void_p(dr.data_ptr()), 1), "mc_batch")
    torch.cuda.synchronize()
    dev = dr.cpu().numpy().view(H.MC_RES_DTYPE)
    assert dev.tobytes() == host.tobytes()


def test_bad_batches_are_refused(hip):
    p = hip.params_from_scene(H.scene_preset("Integer Power", 32, 24))
    t = view_tasks(p)[:4].copy()
    t["kind"][2] = 3
    r = np.zeros(4, H.MC_RES_DTYPE)
    st = H.load().mb3d_hip_mc_batch(C.byref(p), H._p(t), 4, 0, H._p(r), 0)
    assert st == H.ERR_INVALID and "kind" in H.last_error()
    assert len(hip.mc_batch(p, np.zeros(0, H.MC_TASK_DTYPE))) == 0


def test_the_frame_batches_of_the_benchmark(hip):
    """The benchmark's two batches at a small size: every pixel's primary
    ray with the inside probe (mc_primary_tasks), then shadow rays towards
    a light from the surfaces they met (mc_shadow_tasks)."""
    p = hip.params_from_scene(H.scene_bulb(192, 108, seed=1))
    prim = hip.mc_primary_tasks(p, seed=1)
    got, want = run_both(hip, p, prim)
    m = compare(got, want, "bench primary")
    assert m["surfaces"] > 2000
    shad, surf = hip.mc_shadow_tasks(p, want, prim["dir"], (0.6, -0.5, -0.6))
    assert len(shad) == (want["kind"] == 1).sum() and np.all(shad["zzplus"] > -1.0) and np.all(shad["zzplus"] < 1.0)
    sg, sw = run_both(hip, p, shad)
    compare(sg, sw, "bench shadow")
    assert 0.02 < (sw["len"] > 0.01).mean() < 0.98


def test_a_task_of_no_kind_in_device_memory(hip):
    """Device tasks are not checked on the host: a kind out of 0..2 returns
    kind -1 with its seed and no march; the other tasks are unaffected."""
    import torch

    p = hip.params_from_scene(H.scene_preset("Integer Power", 48, 36))
    t = view_tasks(p)
    host = hip.mc_batch(p, t)
    bad = t.copy()
    bad["kind"][::5] = 7
    dt = torch.from_numpy(bad.view(np.uint8).copy()).cuda()
    dr = torch.zeros(len(t) * H.MC_RES_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    H.check(H.load().mb3d_hip_mc_batch(C.byref(p), C.c_void_p(dt.data_ptr()), len(t), 0,
                                       C.c_void_p(dr.data_ptr()), 1), "mc_batch")
    torch.cuda.synchronize()
    dev = dr.cpu().numpy().view(H.MC_RES_DTYPE)
    assert np.all(dev["kind"][::5] == -1) and np.array_equal(dev["seed"][::5], t["seed"][::5])
    keep = np.ones(len(t), bool)
    keep[::5] = False
    assert dev[keep].tobytes() == host[keep].tobytes()
