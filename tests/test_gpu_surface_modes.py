"""The general surface modes on the device (k_normals_ext): smooth normals
1..8 with the roughness, normals on the smooth iteration count, colour
options 1..5 and colour on an iteration (surface_found and its callees,
src/calc.rs:1806-2090, 2471-2517), against the oracle.  The reference's own
scene fixture, tests/data/mandelbulb.m3s (smooth normals 3), is held to the
plain bar and to the reference's tolerances between its card and its CPU
(src/gpu.rs:4240-4243)."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity
from test_gpu_batch import device_out, host

pytestmark = pytest.mark.gpu


def words_equal(g, r, surf):
    """The share of surface pixels whose five G-buffer words are equal, and
    of those whose normal words differ by at most one unit."""
    a, b = g["gbuf"][surf].astype(np.int64), r["gbuf"][surf].astype(np.int64)
    eq = (a == b).all(1).mean()
    na = np.stack([a[:, 0] & 0xFFFF, a[:, 0] >> 16, a[:, 1] & 0xFFFF], 1).astype(np.uint16).view(np.int16)
    nb = np.stack([b[:, 0] & 0xFFFF, b[:, 0] >> 16, b[:, 1] & 0xFFFF], 1).astype(np.uint16).view(np.int16)
    near = (np.abs(na.astype(int) - nb.astype(int)) <= 1).all(1).mean()
    return float(eq), float(near)


def test_the_reference_fixture_scene(hip):
    sc = H.scene_m3s_mandelbulb(400, 300)
    p = hip.params_from_scene(sc)
    g = hip.march(p)  # the numeric DE: the automatic precision marches in FP64
    r = O.render(O.params(sc), threads=0)
    m = parity.compare(g, r, p.width, p.height, p.step_width, 2.0)
    assert m["mask_agree"] >= 0.999 and m["depth_ok_frac"] >= 0.999 and m["normal_ok_frac"] >= 0.999, m
    assert m["iters_equal_inner"] >= 0.999, m
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0) & np.isfinite(g["depth64"]) & (g["depth64"] > 0)
    eq, near = words_equal(g, r, surf)
    assert eq >= 0.99 and near >= 0.999, (eq, near)
    # the reference's card-vs-CPU bar: mean depth-code difference < 20 / 32768,
    # mean normal cosine > 0.97, < 3 % of the normals off by more than 25 degrees
    zg, zr = (g["gbuf"][surf, 2] >> 16).astype(float), (r["gbuf"][surf, 2] >> 16).astype(float)
    assert np.abs(zg - zr).mean() < 20.0
    cosang = (g["normals"][surf] * r["normals"][surf]).sum(1)
    assert cosang.mean() > 0.97 and (cosang < np.cos(np.radians(25))).mean() < 0.03
    assert (g["gbuf"][surf, 2] & 0xFF).any()  # the roughness byte


MODES = [  # (smooth normals, normals on the DE, colour option, colour on iteration)
    (1, 1, 0, 0), (3, 1, 0, 0), (8, 1, 0, 0),
    (0, 1, 1, 0), (0, 1, 2, 0), (0, 1, 3, 0), (0, 1, 4, 0), (0, 1, 5, 0),
    (0, 1, 2, 1), (0, 1, 1, 4), (3, 1, 5, 3),
]


@pytest.mark.parametrize("mode", MODES)
def test_surface_modes_match_the_oracle(hip, mode):
    sm, on_de, col, col_it = mode
    sc = H.scene_bulb(160, 120, seed=3)  # the analytic power-8 DE
    sc.smooth_normals, sc.normals_on_de, sc.color_option, sc.color_on_it = sm, on_de, col, col_it
    p = hip.params_from_scene(sc)
    g = hip.march(p, opts=H.options(precision=H.PREC_F64))
    r = O.render(O.params(sc), threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    assert np.array_equal(np.isfinite(g["depth64"]), np.isfinite(r["depth"]))
    eq, near = words_equal(g, r, surf)
    assert eq >= 0.995 and near >= 0.999, (mode, eq, near)


@pytest.mark.parametrize("sm", [0, 3])
def test_normals_on_the_smooth_iteration_match_the_oracle(hip, sm):
    sc = H.scene_preset("Integer Power", 160, 120)  # the numeric DE: normals on the DE may be off
    sc.first_step_random = 0
    sc.normals_on_de = 0
    sc.smooth_normals = sm
    p = hip.params_from_scene(sc)
    assert not p.normals_on_de
    g = hip.march(p)
    r = O.render(O.params(sc), threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0)
    eq, near = words_equal(g, r, surf)
    assert eq >= 0.995 and near >= 0.999, (sm, eq, near)


def test_fp32_march_with_smooth_normals_meets_the_bar(hip):
    sc = H.scene_bulb(320, 180, seed=5)
    sc.smooth_normals = 3
    sc.color_option = 2
    p = hip.params_from_scene(sc)
    g = hip.march(p)  # the FP32 march, the f64 finish, the general normals pass
    r = O.render(O.params(sc), threads=0)
    m = parity.compare(g, r, p.width, p.height, p.step_width, 2.0)
    assert m["mask_agree"] >= 0.999 and m["depth_ok_frac"] >= 0.999 and m["normal_ok_frac"] >= 0.999, m


def test_batch_with_smooth_normals_equals_single_frames(hip):
    scenes = H.animation(4, seed=9, sizes=((96, 72), (160, 120)))
    for s in scenes:
        s.smooth_normals = 3
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    bufs, out = device_out(n)
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), None), "batch")
    got = host(bufs)
    off = 0
    for i, s in enumerate(scenes):
        m = s.width * s.height
        one = hip.march(arr[i])
        assert np.array_equal(got["gbuf"][off * 5:(off + m) * 5].view(np.uint32), one["gbuf"].ravel()), i
        off += m


def test_ray_batches_refuse_the_general_modes(hip):
    sc = H.scene_bulb(64, 48, seed=1)
    sc.smooth_normals = 3
    p = hip.params_from_scene(sc)
    t = H.mc_primary_tasks(p, seed=1)
    with pytest.raises(RuntimeError, match="ray batches"):
        H.mc_batch(p, t)
