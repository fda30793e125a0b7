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


def compare(a, b, what):
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
    print(what, m, flush=True)
    assert m["flags"] >= 0.999, m
    for k in ("zz2_ok", "zcode", "steps", "normals"):
        assert m[k] >= 0.999, (k, m)
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
    H.check(H.load().mb3d_hip_march_rays(C.byref(p), C.c_void_p(dt.data_ptr()), len(t), 0,
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
    st = H.load().mb3d_hip_march_rays(C.byref(p), H._p(t), 4, 0, H._p(h), 0)
    assert st == H.ERR_INVALID and "pixel" in H.last_error()
    assert len(hip.march_rays(p, np.zeros(0, H.RAY_TASK_DTYPE))) == 0
