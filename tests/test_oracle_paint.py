"""The oracle's painter (oracle/paint_oracle.cpp) pinned on the CPU, and
the product's host-side set-up (mb3d_hip_paint_setup) checked against it.

What pins the painter, beside the reference's own lighting tests
(src/lighting.rs:2228-2257, restated below):
- the diffuse tables against closed forms: the cubic B-spline of
  MakeSplineCoeff reproduces a linear table exactly and adds h^2 f''/6 to a
  quadratic one, which fixes the table layout, the spline weights and the
  index offset (ip = trunc(t) - 1);
- the shading of synthetic G-buffer records against closed forms in which
  one term is left (diffuse with the half-cosine function, specular of a
  mirror normal, the ambient top/bottom blend): they fix the record
  layout, the light direction's sign, the colour scales and the exponent;
- invariants: the background is the depth colour blend from top to
  bottom, hard shadows and ambient occlusion darken, gamma 32 is neutral,
  the result does not depend on the thread count.
"""
import ctypes as C
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O


def scene(w=64, h=48, name="Integer Power"):
    return H.scene_preset(name, w, h)


def rec(n, normal=(0.0, 0.0, -1.0), zpos=20000, rough=0, shadow=0, amb=0, grad=16000, otrap=0):
    """n G-buffer records (the card's 5 words): unit normal in the view
    frame, depth word, roughness byte, shadow word, AO, gradient, trap."""
    nv = np.asarray(normal, np.float64)
    nv = nv / np.linalg.norm(nv, axis=-1, keepdims=True)
    q = np.round(nv * 32767.0).astype(np.int64) & 0xFFFF
    q = np.broadcast_to(q, (n, 3))
    g = np.zeros((n, 5), np.uint32)
    g[:, 0] = q[:, 0] | (q[:, 1] << 16)
    g[:, 1] = q[:, 2] | (np.uint32(shadow) << 16)
    g[:, 2] = (np.uint32(zpos) << 16) | np.uint32(rough)
    g[:, 3] = np.uint32(grad) | (np.uint32(otrap) << 16)
    g[:, 4] = np.uint32(amb)
    return g


def plain_lighting(diffuse_rgb=(200, 150, 100), specular_rgb=(0, 0, 0)):
    """One white light from the viewer (angles 0), no ambient light, no
    fog, the palette one colour."""
    lt = O.lighting_default()
    for i in range(6):
        lt.lights[i].on = 0
    l0 = lt.lights[0]
    l0.on, l0.x_angle, l0.y_angle, l0.diff_func, l0.spec_func = 1, 0.0, 0.0, 2, 4
    l0.color[:] = [255, 255, 255]
    lt.ambient = 0.0
    lt.depth_fog = 0.0
    lt.dyn_fog = 53.0
    for i in range(10):
        lt.palette[i].diffuse[:] = diffuse_rgb
        lt.palette[i].specular[:] = specular_rgb
    return lt


def paint_records(g, lt, w, h, fmt=H.PAINT_U8, sc=None):
    sc = sc or scene(w, h)
    p = H.params_from_scene(sc)
    cam, lv = O.paint_setup(sc, p, lt)
    return O.paint(cam, lv, g, fmt=fmt), cam, lv


# ---- the reference's tests (src/lighting.rs:2228-2257)


def test_cos_tab_monotonic():
    a, b, c = O.cos_tab_val(0, 1.0), O.cos_tab_val(0, 0.5), O.cos_tab_val(0, -0.5)
    assert a > b > c
    assert abs(a - 1.0) < 0.1


def test_default_colors():
    for lt in (O.lighting_default(), H.lighting_default()):
        assert list(lt.palette[0].diffuse) == [0xE1, 0xA0, 0x59]
        assert list(lt.lights[0].color) == [0xFF, 0xE8, 0xA0]


def test_the_far_fog_offset_sets_the_depth_change():
    """LightVals::new(l, 0.1, 0, 4.0, 0.01, 96, 1.0): s_shad_zmul scales by
    (128 - sqrt(255) 11.313708) / (128 - sqrt(128) 11.313708)."""
    sc = scene(96, 72)
    sc.z_step_div = 0.1
    sc.z_end = sc.z_start + 4.0
    p = H.params_from_scene(sc)
    lt = O.lighting_default()
    neutral = O.paint_setup(sc, p, lt)[1].s_shad_zmul
    lt.fog_far_offset = 255.0
    v = O.paint_setup(sc, p, lt)[1].s_shad_zmul
    f = np.float32
    k = (f(128.0) - np.sqrt(f(255.0)) * f(11.313708)) / (f(128.0) - np.sqrt(f(128.0)) * f(11.313708))
    assert abs(f(v) - f(neutral) * k) <= 1e-4 * max(abs(neutral), 1.0)
    # the value itself: 2.2 / z_step_div * 0.7 / z_range * (128 - sqrt(offset) 11.313708), in f32
    want = f(2.2) / f(0.1) * f(0.7) / f(4.0) * (f(128.0) - np.sqrt(f(255.0)) * f(11.313708))
    assert f(v) == want


# ---- the diffuse tables


def test_the_rough_tables_smooth_and_the_plain_ones_follow_their_functions():
    t = O.cos_tables()
    i = np.arange(128)
    d = 1.0 - (i - 2.0) / 60.0
    assert np.allclose(t[2], d * 0.5 + 0.5, atol=1e-7)
    assert np.allclose(t[1], np.maximum(d, 0) ** 2, atol=1e-7)
    assert np.allclose(t[3], (d * 0.5 + 0.5) ** 2, atol=1e-7)
    sel = d > 0.15
    assert np.allclose(t[0][sel], (d[sel] - 0.08) * 1.0869565, atol=1e-7)
    # the rough variants: non-increasing along the table (light to dark)
    for k in range(4, 8):
        assert np.all(np.diff(t[k]) <= 1e-6)


@pytest.mark.parametrize("dotp", [-0.8, -0.3, 0.05, 0.37, 0.71, 0.93])
def test_the_spline_reproduces_the_half_cosine_exactly(dotp):
    """Cubic B-splines have linear precision: function 2 (d/2 + 1/2) comes
    back as it is at any light angle."""
    assert abs(O.cos_tab_val(2, dotp) - (0.5 * dotp + 0.5)) < 2e-7


@pytest.mark.parametrize("dotp", [0.2, 0.45, 0.8])
def test_the_spline_adds_h2_fpp_over_6_to_the_quadratics(dotp):
    """Function 1 (d^2) and 3 ((d/2 + 1/2)^2) with the table step h = 1/60:
    the spline adds h^2 f''/6."""
    h = 1.0 / 60.0
    assert abs(O.cos_tab_val(1, dotp) - (dotp * dotp + h * h * 2.0 / 6.0)) < 3e-7
    assert abs(O.cos_tab_val(3, dotp) - ((0.5 * dotp + 0.5) ** 2 + h * h * 0.5 / 6.0)) < 3e-7


def test_roughness_blends_towards_the_rough_table():
    t = O.cos_tables()
    for dotp in (0.3, 0.8):
        a, b = O.cos_tab_val(0, dotp, 0.0), O.cos_tab_val(0, dotp, 1.0)
        assert abs(O.cos_tab_val(0, dotp, 0.25) - (a + 0.25 * (b - a))) < 1e-6
    assert O.cos_tab_val(0, 0.5, 0.0, sqr=True) == pytest.approx(O.cos_tab_val(0, 0.5) ** 2, rel=1e-6)
    assert t.shape == (8, 128)


# ---- shading closed forms on synthetic records


def _centre(w, h):
    # the pixel whose view ray is the view axis: x + 1 = w / 2, y = h / 2
    return h // 2, w // 2 - 1


def test_diffuse_alone_is_colour_times_light_times_the_half_cosine():
    w, h = 64, 48
    angles = np.radians([0.0, 20.0, 45.0, 70.0, 85.0])
    normals = np.stack([np.sin(angles), np.zeros_like(angles), -np.cos(angles)], axis=1)
    g = np.concatenate([rec(1, n) for n in normals] + [rec(w * h - len(angles))])
    lt = plain_lighting()
    rgb, cam, lv = paint_records(g, lt, w, h)
    assert lv.n_lights == 1 and list(lv.lights[0].ln) == pytest.approx([0.0, 0.0, -1.0], abs=1e-7)
    pal = np.array([200, 150, 100], np.float64)
    s_diff = 120 * 0.02
    q = np.round(normals * 32767.0) / 32767.0 * np.float32(3.0518509476e-5) * 32767.0
    dot = -q[:, 2]
    want = np.clip(pal[None, :] / 255.0 * s_diff * 255.0 * (0.5 * dot[:, None] + 0.5), 0, 255)
    got = rgb.reshape(-1, 3)[: len(angles)].astype(np.float64)
    assert np.abs(got - np.round(want)).max() <= 1, (got, want)


@pytest.mark.parametrize("spec_func", [2, 4, 6])
def test_specular_alone_is_the_mirror_lobe(spec_func):
    """Palette diffuse 0, specular colour s: at the centre pixel a normal
    tilted by a from the view axis reflects the viewer's light at 2a, so
    the specular is s/255 * 255 * s_spec * cos(2a)^(2 << spec_func)."""
    w, h = 64, 48
    yc, xc = _centre(w, h)
    lt = plain_lighting(diffuse_rgb=(0, 0, 0), specular_rgb=(180, 120, 60))
    lt.lights[0].spec_func = spec_func
    lt.specular = 100.0  # s_spec = 2
    for a_deg in (0.0, 6.0, 14.0):
        a = math.radians(a_deg)
        n = (math.sin(a), 0.0, -math.cos(a))
        g = rec(w * h, n)
        rgb, _, lv = paint_records(g, lt, w, h)
        nq = np.round(np.array(n) * 32767.0) * np.float32(3.0518509476e-5)
        view = np.array([0.0, 0.0, 1.0])
        r = view - 2.0 * np.dot(nq, view) * nq
        ds = -r[2]
        pf = 2 << spec_func
        want = np.array([180, 120, 60]) / 255.0 * 255.0 * 2.0 * ds ** pf
        assert lv.lights[0].pow_func == pf
        assert np.abs(rgb[yc, xc].astype(np.float64) - np.round(np.clip(want, 0, 255))).max() <= 1, (a_deg, rgb[yc, xc], want)


def test_ambient_blends_top_and_bottom_colours_by_the_normal():
    """No lights, ambient 90 (factor 1): a normal pointing along +y (down
    the picture) takes the bottom colour, -y the top one."""
    w, h = 32, 24
    lt = plain_lighting(diffuse_rgb=(255, 255, 255))
    lt.lights[0].on = 0
    lt.ambient = 90.0
    lt.amb_top[:] = [10, 20, 30]
    lt.amb_bottom[:] = [200, 100, 50]
    for ny, t2 in ((1.0, 1.0), (-1.0, 0.0), (0.0, 0.5)):
        n = (0.0, ny, -1.0) if ny == 0.0 else (0.0, ny, 0.0)
        g = rec(w * h, n)
        rgb, _, _ = paint_records(g, lt, w, h)
        nyq = round(n[1] / math.hypot(*n) * 32767.0) * 3.0518509476e-5
        tt = nyq * 0.5 + 0.5
        want = np.array([10, 20, 30]) * (1 - tt) + np.array([200, 100, 50]) * tt
        assert abs(tt - t2) < 1e-4
        assert np.abs(rgb[5, 5].astype(np.float64) - np.round(want)).max() <= 1, (ny, rgb[5, 5], want)


# ---- invariants


def test_the_background_is_the_depth_colour_blend():
    w, h = 40, 30
    g = np.zeros((w * h, 5), np.uint32)
    g[:, 2] = 32768 << 16
    lt = O.lighting_default()
    rgb, _, _ = paint_records(g, lt, w, h)
    top = np.array(lt.depth_col, np.float64)
    bot = np.array(lt.depth_col2, np.float64)
    y = np.arange(h)[:, None] / h
    want = top[None, :] + y * (bot - top)[None, :]
    assert np.array_equal(rgb[0, 0], np.array(lt.depth_col, np.uint8))
    assert np.abs(rgb[:, 7].astype(np.float64) - want).max() <= 1.0
    assert np.all(rgb == rgb[:, :1])  # the same along a row


def _bulb(w=96, h=72):
    sc = scene(w, h)
    p = H.params_from_scene(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], deao_quality=1, params=p)
    r = O.render(O.params(sc), job, threads=0)
    return sc, p, r


def _with_shadow_bits(g, surf, set_bits=0, clear_bits=0):
    """The G-buffer with the shadow word's bits (word 1, high half) of the
    surface records set and cleared."""
    out = g.copy()
    w1 = out[surf, 1]
    out[surf, 1] = (w1 & ~np.uint32(clear_bits << 16)) | np.uint32(set_bits << 16)
    return out


def test_hard_shadows_darken_and_soft_shadows_scale():
    sc, p, r = _bulb()
    g = r["gbuf"]
    surf = (g[:, 2] >> 16) < 32768
    lt = O.lighting_default()
    lt.hs_calced = 1
    cam, lv = O.paint_setup(sc, p, lt)
    lit = O.paint(cam, lv, _with_shadow_bits(g, surf, clear_bits=0x400))
    dark = O.paint(cam, lv, _with_shadow_bits(g, surf, set_bits=0x400))
    assert np.all(dark.astype(int) <= lit.astype(int))
    assert dark.sum() < lit.sum()
    # a light without a calculated shadow ignores the bit
    lt.hs_calced = 0
    cam, lv0 = O.paint_setup(sc, p, lt)
    assert np.array_equal(O.paint(cam, lv0, _with_shadow_bits(g, surf, set_bits=0x400)),
                          O.paint(cam, lv0, _with_shadow_bits(g, surf, clear_bits=0x400)))
    # soft: the shadow word's top 6 bits scale the light, 63 * (1/63) in
    # f32 being 1 to an ulp: 63 is the lit picture to a unit of the last
    # bit, 0 the light's diffuse and specular parts gone
    lt.hs_calced, lt.hs_soft = 1, 1
    cam, lv = O.paint_setup(sc, p, lt)
    assert lv.lights[0].hs_mask == 0xFFFF
    soft63 = O.paint(cam, lv, _with_shadow_bits(g, surf, set_bits=0xFC00))
    soft0 = O.paint(cam, lv, _with_shadow_bits(g, surf, clear_bits=0xFC00))
    assert np.abs(soft63.astype(int) - lit.astype(int)).max() <= 1
    assert np.abs(soft0.astype(int) - dark.astype(int)).max() <= 1


def test_ambient_occlusion_darkens():
    sc, p, r = _bulb()
    g = r["gbuf"].copy()
    surf = (g[:, 2] >> 16) < 32768
    cam, lv = O.paint_setup(sc, p, O.lighting_default())
    prev = None
    for amb in (0, 4000, 9000, 16383):
        gg = g.copy()
        gg[surf, 4] = amb
        s = O.paint(cam, lv, gg)[surf.reshape(72, 96)].astype(np.int64).sum()
        if prev is not None:
            assert s < prev
        prev = s


def test_gamma_32_is_neutral_and_orders_the_others():
    sc, p, r = _bulb()
    sums = {}
    for gm in (20.0, 32.0, 45.0):
        lt = O.lighting_default()
        lt.gamma = gm
        cam, lv = O.paint_setup(sc, p, lt)
        sums[gm] = O.paint(cam, lv, r["gbuf"]).astype(np.int64).sum()
        if gm == 32.0:
            assert lv.gamma_h == 0
    assert sums[20.0] < sums[32.0] < sums[45.0]


def test_a_visible_light_glows_around_its_direction():
    """A global light straight ahead (x angle pi): on the background its
    glow peaks at the pixel of the view axis and falls off outwards."""
    w, h = 64, 48
    g = np.zeros((w * h, 5), np.uint32)
    g[:, 2] = 32768 << 16
    lt = O.lighting_default()
    lt.lights[0].x_angle, lt.lights[0].y_angle, lt.lights[0].visible = math.pi, 0.0, 2
    on = paint_records(g, lt, w, h)[0].astype(np.int64)
    lt.lights[0].visible = 0
    off = paint_records(g, lt, w, h)[0].astype(np.int64)
    glow = (on - off).sum(axis=2)
    yc, xc = _centre(w, h)
    assert glow[yc, xc] == glow.max() > 0
    assert glow[yc, xc] > glow[yc, xc + 1] > glow[yc, xc + 3] == 0
    assert glow[yc, xc - 1] == glow[yc, xc + 1] and glow[yc - 1, xc] == glow[yc + 1, xc]


def test_the_painter_does_not_depend_on_the_thread_count():
    sc, p, r = _bulb()
    lt = O.lighting_default()
    lt.lights[1].on = 1
    lt.lights[1].positional, lt.lights[1].visible = 1, 4
    lt.lights[1].position[:] = [-0.9, 0.3, -1.1]
    cam, lv = O.paint_setup(sc, p, lt)
    a = O.paint(cam, lv, r["gbuf"], threads=1)
    b = O.paint(cam, lv, r["gbuf"], threads=5)
    assert np.array_equal(a, b)


def test_sixteen_bit_and_float_agree_with_eight_bit():
    sc, p, r = _bulb(48, 36)
    cam, lv = O.paint_setup(sc, p, O.lighting_default())
    a = O.paint(cam, lv, r["gbuf"], fmt=H.PAINT_U8).astype(np.float64)
    b = O.paint(cam, lv, r["gbuf"], fmt=H.PAINT_U16).astype(np.float64) / 257.0
    c = O.paint(cam, lv, r["gbuf"], fmt=H.PAINT_F32).astype(np.float64) * 255.0
    assert np.abs(a - b).max() <= 0.5 + 1e-9
    assert np.abs(np.clip(c, 0, 255) - b).max() <= 0.5 / 257.0 * 255.0 + 1e-3


# ---- the product's set-up against the oracle's (host code, no GPU)


def _variants():
    lt = O.lighting_default()
    yield "default", lt
    lt = O.lighting_default()
    lt.internal_gamma2, lt.ex_mode, lt.gamma, lt.roughness = 1, 1, 41.0, 120.0
    lt.fine_col_adj_on, lt.fine_col_adj[:] = 1, [12, 47]
    yield "gamma2-exmode", lt
    lt = O.lighting_default()
    for i, (pos, vis) in enumerate(((1, 2), (1, 4), (0, 6), (1, 8), (0, 0))):
        L = lt.lights[i + 1]
        L.on, L.positional, L.visible, L.amplitude = 1, pos, vis, 0.7 + 0.1 * i
        L.position[:] = [-0.9 + 0.3 * i, 0.3 - 0.1 * i, -1.1 + 0.2 * i]
        L.x_angle, L.y_angle, L.spec_func, L.diff_func = 0.3 * i, -0.2 * i, i % 8, i % 4
        L.relative_to_object = i % 2
    lt.hs_calced, lt.hs_set_cos, lt.hs_soft = 0b101011, 1, 0
    lt.vol_light = 2
    lt.dyn_fog, lt.dfog_options, lt.far_fog, lt.depth_func = 61.0, 3, 1, 2
    yield "many-lights", lt
    lt = O.lighting_default()
    lt.amb_rel_obj, lt.has_amb_frame = 1, 1
    lt.amb_frame[0][:] = [0.0, 1.0, 0.0]
    lt.amb_frame[1][:] = [0.0, 0.0, 1.0]
    lt.amb_frame[2][:] = [1.0, 0.0, 0.0]
    lt.amb_fov = 0.8
    lt.color_cycling, lt.color_on_otrap, lt.no_col_ipol, lt.var_col_z = 1, 1, 1, 30.0
    lt.palette_alpha_set = 1
    lt.palette_alpha[:] = list(range(10, 110, 10))
    yield "amb-frame-cycling", lt


@pytest.mark.parametrize("name", ["default", "gamma2-exmode", "many-lights", "amb-frame-cycling"])
def test_product_setup_equals_the_oracle(name):
    lt = dict(_variants())[name]
    for sname in ("Integer Power", "Amazing Box"):
        sc = scene(120, 90, sname)
        sc.dfog_on_it = 1 if name == "many-lights" else 0
        p = H.params_from_scene(sc)
        ch, lh = H.paint_setup(sc, p, lt)
        co, lo = O.paint_setup(sc, p, lt)
        assert bytes(ch) == bytes(co)
        if bytes(lh) != bytes(lo):
            dh, do = H.struct_to_dict(lh), H.struct_to_dict(lo)
            bad = {k: (dh[k], do[k]) for k in dh if dh[k] != do[k]}
            raise AssertionError(bad)


def test_setup_refuses_what_the_painter_lacks():
    sc = scene()
    p = H.params_from_scene(sc)
    for field, msg in (("bg_image", "background"), ("diff_map", "diffuse"), ("bg_ambient", "ambient")):
        lt = H.lighting_default()
        setattr(lt, field, 1)
        with pytest.raises(RuntimeError, match=msg):
            H.paint_setup(sc, p, lt)
    lt = H.lighting_default()
    lt.lights[2].map = 3
    with pytest.raises(RuntimeError, match="light maps"):
        H.paint_setup(sc, p, lt)


def test_lighting_defaults_agree():
    assert bytes(H.lighting_default()) == bytes(O.lighting_default())
    assert H.lighting_default().vol_light == -1
    assert C.sizeof(H.Lighting) == C.sizeof(O.lighting_default())


# ---- the G-buffer's file records (gbuffer::pack / unpack)


def test_m3i_round_trip_record():
    """src/m3p.rs:1671-1692: 24 x 16 default records, record 5 with normal
    (1, -2, 3) and gradient 1234: its first 6 bytes are 1 0 254 255 3 0,
    and the gradient comes back."""
    g = np.zeros((24 * 16, 5), np.uint32)
    g[:, 2] = 32768 << 16  # SiLight::default: background
    g[5, 0] = 1 | ((-2 & 0xFFFF) << 16)
    g[5, 1] = 3
    g[5, 3] = 1234
    d = O.gbuf_pack(g)
    assert d.size == 24 * 16 * 18
    assert list(d[5 * 18:5 * 18 + 6]) == [1, 0, 254, 255, 3, 0]
    back = O.gbuf_unpack(d)
    assert back[5, 3] & 0xFFFF == 1234
    assert np.array_equal(back, g)


def test_pack_is_the_little_endian_field_order():
    """One record by hand: normal, zpos_fine, shadow, amb_shadow,
    si_gradient, otrap, little endian (gbuffer::pack)."""
    g = np.array([[0x2222_1111, 0xABCD_3333, 0x4433_2211, 0x7788_5566, 0x0001_0FED]], np.uint32)
    d = O.gbuf_pack(g)
    want = [0x11, 0x11, 0x22, 0x22, 0x33, 0x33, 0x11, 0x22, 0x33, 0x44, 0xCD, 0xAB, 0xED, 0x0F, 0x66, 0x55, 0x88, 0x77]
    assert list(d) == want
    # the card's overflow flag (bit 16 of the last word) is not a file field
    assert O.gbuf_unpack(d)[0, 4] == 0x0FED


def test_pack_round_trips_a_marched_gbuffer():
    sc, p, r = _bulb(48, 36)
    g = r["gbuf"]
    assert np.array_equal(O.gbuf_unpack(O.gbuf_pack(g)), g & np.array([0xFFFFFFFF] * 4 + [0xFFFF], np.uint32))
