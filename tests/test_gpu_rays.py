"""Secondary rays on the device (mb3d_hip_march_rays, k_rays: the march of
the reference's reflection batches, SrTracer::march_ray,
src/reflect.rs:1464-1674) against the oracle on the same tasks.

Both sides march in f64 in the reference's operation order; they differ
by the device's libm in the pixel's view ray (an ulp of zzplus) and in the
formulas that call sin/cos/pow.  The bar per batch: hit flags equal for
>= 99.9 % of the rays; over the rays both hit, the ray length within 1e-4
steps, the depth code and the step count equal, and the normal within
0.5 degrees, each for >= 99.9 %."""
import ctypes as C

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def normals_of(g):
    v = np.stack([g[:, 0] & 0xFFFF, g[:, 0] >> 16, g[:, 1] & 0xFFFF], axis=1).astype(np.int64)
    v = np.where(v >= 32768, v - 65536, v).astype(np.float64)
    return v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-300)


def compare(a, b, what, post=False):
    fa, fb = a["flags"], b["flags"]
    agree = (fa == fb).mean()
    both = ((fa & H.RAY_HIT) != 0) & ((fb & H.RAY_HIT) != 0)
    n = int(both.sum())
    dz = np.abs(a["zz2"][both] - b["zz2"][both])
    ga, gb = a["gbuf"][both], b["gbuf"][both]
    zcode = (ga[:, 2] == gb[:, 2]).mean() if n else 1.0
    steps = ((ga[:, 1] >> 16) == (gb[:, 1] >> 16)).mean() if n else 1.0
    cos = np.sum(normals_of(ga) * normals_of(gb), axis=1)
    nok = (cos > np.cos(np.radians(0.5))).mean() if n else 1.0
    m = {"rays": len(fa), "hits": n, "flags": agree, "zz2_ok": (dz <= 1e-4).mean() if n else 1.0, "zcode": zcode,
         "steps": steps, "normals": nok, "grad": (ga[:, 3] == gb[:, 3]).mean() if n else 1.0}
    if post and n:
        m["shadow_bits"] = (((ga[:, 1] >> 26) == (gb[:, 1] >> 26))).mean()
        m["hard_bits"] = ((((ga[:, 1] >> 16) & 0xFC00) == ((gb[:, 1] >> 16) & 0xFC00))).mean()
        da = np.abs((ga[:, 4] & 0xFFFF).astype(np.int64) - (gb[:, 4] & 0xFFFF).astype(np.int64))
        m["ao_exact"] = (da == 0).mean()
        m["ao_mean_diff"] = da.mean() / 16383.0
    print(what, m, flush=True)
    assert m["flags"] >= 0.999, m
    for k in ("zz2_ok", "zcode", "steps", "normals"):
        assert m[k] >= 0.999, (k, m)
    if post and n:
        assert m["hard_bits"] >= 0.999 and m["ao_exact"] >= 0.99 and m["ao_mean_diff"] < 1e-3, m
    return m


def reflection_case(hip, sc, depth_sqrt=False):
    p = hip.params_from_scene(sc)
    g = hip.march(p)
    t = hip.reflection_tasks(p, g)
    assert len(t) > 1000
    got = hip.march_rays(p, t, depth_sqrt)
    want = O.march_rays(O.params(sc), t, depth_sqrt)
    return p, t, got, want


@pytest.mark.parametrize("preset", ["Integer Power", "Amazing Box", "Quaternion", "Bulbox"])
def test_reflection_rays_match_the_oracle(hip, preset):
    sc = H.scene_preset(preset, 320, 240)
    p, t, got, want = reflection_case(hip, sc)
    m = compare(got, want, preset)
    assert m["hits"] > 50
    c = H.counters()
    assert c["pixels"] == len(t) and c["de_evals_f64"] > 0


def test_the_analytic_bulb_and_the_far_fog_depth(hip):
    """The benchmark's power-8 bulb (analytic DE, the FP32 kinds' f64 path)
    with the square-root depth code."""
    sc = H.scene_bulb(320, 240, seed=2)
    p, t, got, want = reflection_case(hip, sc, depth_sqrt=True)
    compare(got, want, "bulb analytic, sqrt depth")


def test_view_rays_inside_and_cut_rays(hip):
    """Every pixel's own view ray from its start plane, a third of them
    flagged inside (no march) and a third cut at 40 steps."""
    sc = H.scene_preset("Amazing Box", 200, 150)
    p = hip.params_from_scene(sc)
    rays, step = hip.rays(p)
    n = p.width * p.height
    t = np.zeros(n, H.RAY_TASK_DTYPE)
    t["origin"] = rays[:, :3]
    t["dir"] = (rays[:, 3:] * step[:, None]).astype(np.float32)
    t["pixel"] = np.arange(n, dtype=np.uint32)
    t["inside"][::3] = 1
    t["max_len"][1::3] = 40.0
    got = hip.march_rays(p, t)
    want = O.march_rays(O.params(sc), t)
    compare(got, want, "view rays")
    assert not (got["flags"][::3] & H.RAY_HIT).any()
    assert np.all(got["zz2"][::3] == p.zend)


def test_rays_on_the_device(hip):
    """Tasks and hits in device memory (the reflection pass keeps them
    there between bounces)."""
    import torch

    sc = H.scene_preset("Integer Power", 240, 180)
    p = hip.params_from_scene(sc)
    g = hip.march(p)
    t = hip.reflection_tasks(p, g)
    host = hip.march_rays(p, t)
    dt = torch.from_numpy(t.view(np.uint8).copy()).cuda()
    dh = torch.zeros(len(t) * H.RAY_HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    H.check(H.load().mb3d_hip_march_rays(C.byref(p), C.c_void_p(dt.data_ptr()), len(t), 0, None,
                                         C.c_void_p(dh.data_ptr()), 1), "march_rays")
    torch.cuda.synchronize()
    dev = dh.cpu().numpy().view(H.RAY_HIT_DTYPE)
    assert dev.tobytes() == host.tobytes()


def test_bad_ray_batches_are_refused(hip):
    sc = H.scene_preset("Integer Power", 32, 24)
    p = hip.params_from_scene(sc)
    t = np.zeros(4, H.RAY_TASK_DTYPE)
    t["pixel"] = [0, 1, 2, 32 * 24]
    h = np.zeros(4, H.RAY_HIT_DTYPE)
    st = H.load().mb3d_hip_march_rays(C.byref(p), H._p(t), 4, 0, None, H._p(h), 0)
    assert st == H.ERR_INVALID and "pixel" in H.last_error()
    t["pixel"][3] = 3
    o = H.ray_opts(p, ao_max_len=1.0)
    o.ao_first_random = 1
    st = H.load().mb3d_hip_march_rays(C.byref(p), H._p(t), 4, 0, C.byref(o), H._p(h), 0)
    assert st == H.UNSUPPORTED and "random" in H.last_error()
    o = H.ray_opts(p, soft_radius=1.0)
    st = H.load().mb3d_hip_march_rays(C.byref(p), H._p(t), 4, 0, C.byref(o), H._p(h), 0)
    assert st == H.ERR_INVALID and "soft" in H.last_error()
    assert len(hip.march_rays(p, np.zeros(0, H.RAY_TASK_DTYPE))) == 0


@pytest.mark.parametrize("case", ["hard-global-positional", "soft-positional", "ao-all-qualities"])
def test_the_hit_pass_matches_the_oracle(hip, case):
    """calc_hs and deao_1pos at the hits (k_ray_post) against the oracle:
    two hard shadows (a global light and a positional one), one soft
    shadow of a positional light, and DE AO of every quality."""
    sc = H.scene_preset("Integer Power", 320, 240)
    p = hip.params_from_scene(sc)
    g = hip.march(p)
    t = hip.reflection_tasks(p, g)
    t["ao_q"] = np.arange(len(t)) % 4
    if case == "hard-global-positional":
        o = H.ray_opts(p, lights=[(0.6, -0.5, -0.6)], positional=[((-0.9, 0.3, -1.1), 40.0)], hs_mul=1.0)
    elif case == "soft-positional":
        o = H.ray_opts(p, positional=[((0.8, -1.2, -0.9), 40.0)], soft_radius=2.0, hs_mul=1.0)
        o.light_visible[0] = 2
    else:
        o = H.ray_opts(p, lights=[(0.2, -0.9, -0.3)], ao_max_len=1.0)
    got = hip.march_rays(p, t, opts=o)
    want = O.march_rays(O.params(sc), t, opts=o)
    m = compare(got, want, case, post=True)
    if case == "ao-all-qualities":
        assert (got["gbuf"][(got["flags"] & H.RAY_HIT) != 0, 4] != 5000).mean() > 0.9
    assert m["hits"] > 1000
