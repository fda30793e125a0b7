"""The HIP painter (mb3d_hip_paint) against the oracle's painter on the
same G-buffer records: the reference's f32 shading (LightVals::
pixel_color_s, src/lighting.rs:1739-2150) restated in the same operation
order, so the colours agree exactly but where the device's f32 sine,
cosine or arcsine of a view ray rounds differently from glibc's (rare
last-bit ties): the bar is >= 99.99 % of the channels equal and none off
by more than one unit (8 and 16 bit), 2e-6 relative in float.

Records: random ones that reach every branch (background, inside
colours, the ground, hard and soft shadow bits, saturated ambient
occlusion, roughness), and marched ones (the march's G-buffer left on the
device and painted there)."""
import ctypes as C
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def random_records(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    q = (np.round(v * 32767.0).astype(np.int64)) & 0xFFFF
    zpos = rng.integers(0, 40000, n)          # about 1 in 5 background (>= 32768)
    rough = rng.integers(0, 256, n)
    steps = rng.integers(0, 1024, n)
    shadow_bits = rng.integers(0, 64, n)
    shadow = steps | (shadow_bits << 10)
    amb = rng.integers(0, 17000, n)
    grad = rng.integers(0, 65536, n)          # half of them inside colours (bit 15)
    otrap = rng.integers(0, 65536, n)
    otrap[rng.random(n) < 0.05] = 0xFFFF      # ground records
    g = np.zeros((n, 5), np.uint32)
    g[:, 0] = q[:, 0] | (q[:, 1] << 16)
    g[:, 1] = q[:, 2] | (shadow << 16)
    g[:, 2] = (zpos << 16) | rough
    g[:, 3] = grad | (otrap << 16)
    g[:, 4] = amb
    return g


def variants():
    out = {}
    out["default"] = O.lighting_default()
    lt = O.lighting_default()
    lt.internal_gamma2, lt.ex_mode, lt.gamma, lt.roughness = 1, 1, 41.0, 120.0
    lt.fine_col_adj_on, lt.fine_col_adj[:] = 1, [12, 47]
    lt.hs_calced = 0b11
    lt.lights[1].on = 1
    out["gamma2-exmode"] = lt
    lt = O.lighting_default()
    for i, (pos, vis) in enumerate(((1, 2), (1, 4), (0, 6), (1, 8), (0, 0))):
        L = lt.lights[i + 1]
        L.on, L.positional, L.visible, L.amplitude = 1, pos, vis, 0.7 + 0.1 * i
        L.position[:] = [-0.9 + 0.3 * i, 0.3 - 0.1 * i, -1.1 + 0.2 * i]
        L.x_angle, L.y_angle, L.spec_func, L.diff_func = 0.3 * i + 2.8, -0.2 * i, i % 8, i % 4
        L.relative_to_object = i % 2
    lt.hs_calced, lt.hs_set_cos = 0b101011, 1
    lt.vol_light = 2
    lt.dyn_fog, lt.dfog_options, lt.far_fog, lt.depth_func = 61.0, 3, 1, 2
    out["many-lights"] = lt
    lt = O.lighting_default()
    lt.amb_rel_obj, lt.has_amb_frame = 1, 1
    lt.amb_frame[0][:] = [0.0, 1.0, 0.0]
    lt.amb_frame[1][:] = [0.0, 0.0, 1.0]
    lt.amb_frame[2][:] = [1.0, 0.0, 0.0]
    lt.amb_fov = 0.8
    lt.color_cycling, lt.color_on_otrap, lt.no_col_ipol, lt.var_col_z = 1, 1, 1, 30.0
    lt.hs_calced, lt.hs_soft = 1, 1
    lt.dyn_fog, lt.dfog_options, lt.gamma = 48.0, 1, 22.0
    lt.interior_start, lt.interior_end = 40.0, 75.0
    out["amb-frame-cycling-soft"] = lt
    return out


def check_equal(a, b, fmt):
    if fmt == H.PAINT_F32:
        rel = np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b.astype(np.float64)), 1e-6)
        assert rel.max() <= 2e-6, rel.max()
        return
    d = np.abs(a.astype(np.int64) - b.astype(np.int64))
    print("channels differing:", int((d != 0).sum()), "of", d.size, flush=True)
    assert d.max() <= 1 and (d == 0).mean() >= 0.9999, (d.max(), (d != 0).sum())


@pytest.mark.parametrize("fmt", [H.PAINT_U8, H.PAINT_U16, H.PAINT_F32])
@pytest.mark.parametrize("name", ["default", "gamma2-exmode", "many-lights", "amb-frame-cycling-soft"])
def test_random_records_match_the_oracle(hip, name, fmt):
    w, h = 320, 200
    sc = H.scene_preset("Integer Power", w, h)
    p = hip.params_from_scene(sc)
    lt = variants()[name]
    cam, lv = H.paint_setup(sc, p, lt)
    lv.has_ground = 1
    lv.ground_dif[:] = [0.3, 0.5, 0.2]
    lv.ground_spe[:] = [0.6, 0.6, 0.6]
    g = random_records(w * h, seed={"default": 1, "gamma2-exmode": 2, "many-lights": 3}.get(name, 4) + 10 * fmt)
    got = H.paint(cam, lv, g, fmt=fmt)
    want = O.paint(cam, lv, g, fmt=fmt)
    print(name, fmt, got.reshape(-1, 3).mean(axis=0), flush=True)
    check_equal(got, want, fmt)


def test_a_rectangle_of_the_picture(hip):
    """Rows and columns offset: the pixel position enters the view ray,
    the depth colour and the light's geometry."""
    sc = H.scene_preset("Integer Power", 400, 300)
    p = hip.params_from_scene(sc)
    lt = variants()["many-lights"]
    cam, lv = H.paint_setup(sc, p, lt)
    x0, y0, w, h = 37, 91, 211, 57
    g = random_records(w * h, seed=5)
    got = H.paint(cam, lv, g, x0, y0, w, h)
    want = O.paint(cam, lv, g, x0, y0, w, h)
    check_equal(got, want, H.PAINT_U8)
    assert got.shape == (h, w, 3)


@pytest.mark.parametrize("preset", ["Integer Power", "Amazing Box"])
def test_marched_pictures_painted_on_the_device(hip, preset):
    """March and post on the device, paint the G-buffer where the march left
    it (no host round trip), and compare with the oracle's painter on the
    same records -- then the march's own parity: the oracle's marched and
    painted picture against the device's."""
    import torch

    w, h = 480, 360
    sc = H.scene_preset(preset, w, h)
    p = hip.params_from_scene(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], deao_quality=1, params=p)
    n = w * h
    gbuf = torch.zeros(n * 5, dtype=torch.int32, device="cuda")
    out = H.Out(None, None, None, None, gbuf.data_ptr(), 1, 0)
    lib = H.load()
    H.check(lib.mb3d_hip_march(C.byref(p), C.byref(job), 0, h, C.byref(out), None, H.PROGRESS_FN(0),
                               H.CANCEL_FN(0), None), "march")
    lt = O.lighting_default()
    lt.hs_calced = 1
    cam, lv = H.paint_setup(sc, p, lt)
    rgb = torch.zeros(n * 3, dtype=torch.uint8, device="cuda")
    H.check(lib.mb3d_hip_paint(C.byref(cam), C.byref(lv), C.c_void_p(gbuf.data_ptr()), 0, 0, w, h, H.PAINT_U8,
                               C.c_void_p(rgb.data_ptr()), 1), "paint")
    torch.cuda.synchronize()
    g_host = gbuf.cpu().numpy().view(np.uint32).reshape(n, 5)
    got = rgb.cpu().numpy().reshape(h, w, 3)
    check_equal(got, O.paint(cam, lv, g_host), H.PAINT_U8)
    assert np.array_equal(got, H.paint(cam, lv, g_host))
    # end to end against the oracle's march + paint
    r = O.render(O.params(sc), job, threads=0)
    ref = O.paint(*O.paint_setup(sc, p, lt), r["gbuf"])
    d = np.abs(got.astype(np.int64) - ref.astype(np.int64)).max(axis=2)
    print(preset, "pixels within 2:", (d <= 2).mean(), "mean:", d.mean(), flush=True)
    assert (d <= 2).mean() >= 0.99 and d.mean() < 0.5


def test_bad_arguments_are_refused(hip):
    sc = H.scene_preset("Integer Power", 32, 24)
    p = hip.params_from_scene(sc)
    cam, lv = H.paint_setup(sc, p)
    g = random_records(32 * 24, 1)
    out = np.zeros(32 * 24 * 12, np.uint8)
    st = H.load().mb3d_hip_paint(C.byref(cam), C.byref(lv), H._p(g), 0, 0, 32, 24, 7, H._p(out), 0)
    assert st == H.ERR_INVALID and "format" in H.last_error()
    bad = H.LightVals.from_buffer_copy(lv)
    bad.n_lights = 9
    with pytest.raises(RuntimeError, match="light count"):
        H.paint(cam, bad, g)
    bad = H.LightVals.from_buffer_copy(lv)
    bad.lights[0].diff_func = 5
    with pytest.raises(RuntimeError, match="diffuse"):
        H.paint(cam, bad, g)
    assert H.paint(cam, lv, np.zeros((0, 5), np.uint32), w=0, h=0).shape == (0, 0, 3)


def test_gbuf_records_match_the_oracle(hip):
    """mb3d_hip_gbuf_pack / unpack against the oracle's gbuffer::pack /
    unpack on random card records (with the overflow flag set at random),
    from host memory and on the device (also from an odd byte address)."""
    import torch

    n = 100_003
    g = random_records(n, seed=9)
    g[::7, 4] |= 0x10000
    want = O.gbuf_pack(g)
    got = H.gbuf_pack(g)
    assert np.array_equal(got, want)
    back = H.gbuf_unpack(got)
    assert np.array_equal(back, O.gbuf_unpack(want))
    assert np.array_equal(back, g & np.array([0xFFFFFFFF] * 4 + [0xFFFF], np.uint32))
    lib = H.load()
    dg = torch.from_numpy(g.view(np.int32).copy()).cuda()
    raw = torch.zeros(n * 18 + 1, dtype=torch.uint8, device="cuda")
    for off in (0, 1):
        H.check(lib.mb3d_hip_gbuf_pack(H.C.c_void_p(dg.data_ptr()), n, H.C.c_void_p(raw.data_ptr() + off), 1), "pack")
        torch.cuda.synchronize()
        assert np.array_equal(raw.cpu().numpy()[off:off + n * 18], want), off
        dback = torch.zeros(n * 5, dtype=torch.int32, device="cuda")
        H.check(lib.mb3d_hip_gbuf_unpack(H.C.c_void_p(raw.data_ptr() + off), n, H.C.c_void_p(dback.data_ptr()), 1),
                "unpack")
        torch.cuda.synchronize()
        assert np.array_equal(dback.cpu().numpy().view(np.uint32).reshape(n, 5), back), off
