"""The cone rays on the device (k_cone, the cone march of src/cone.rs and
src/calc.rs:2357-2435 with one level of blocks, the queue's tiles) against
the oracle's cone march and the plain march."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity
from test_gpu_batch import device_out, host

pytestmark = pytest.mark.gpu


def scenes():
    # (bulb_wide: rows of 40 tiles of 8 x 8, so k_cone_bg takes background
    # tiles four side by side as well as one at a time)
    return [("bulb", H.scene_bulb(240, 136, seed=4)), ("bulb_numeric_f32", H.scene_bulb(240, 136, seed=5,
                                                                                        de_mode=H.DE_REFERENCE)),
            ("box", H.scene_box(240, 136, seed=3, step_cap=500)), ("hybrid", H.scene_hybrid(240, 136, seed=2)),
            ("bulb_wide", H.scene_bulb(320, 184, seed=8))]


@pytest.mark.parametrize("i", range(5))
@pytest.mark.parametrize("tile", [(8, 8), (16, 8)])
def test_cone_background_leaves_the_plain_picture(hip, i, tile):
    """Mode 1 (a tile whose cone ray passes the end is background): depths,
    iteration counts and normals equal the plain march's on the device; the
    background pixels' step counts are the oracle's cone counts, so the
    device's cone rays found the same empty tiles with the same fog counts."""
    name, sc = scenes()[i]
    p = H.params_from_scene(sc)
    # (the numeric DE's scenes take the FP64 march automatically, which has
    # no cone pass: the FP32 iterate is asked for)
    prec = H.PREC_F32 if name in ("bulb_numeric_f32", "hybrid") else H.PREC_AUTO
    plain = hip.march(p, opts=H.options(precision=prec, tile=tile))
    cone = hip.march(p, opts=H.options(precision=prec, tile=tile, cone=True))
    c = H.counters()
    for k in ("depth", "depth64", "iters", "normals"):
        assert np.array_equal(cone[k], plain[k]), (name, k)
    ref = O.render(O.params(sc), threads=0, cone=(tile[0], tile[1], 1))
    bg = ~np.isfinite(ref["depth"])
    if name != "box":
        assert c["cone_background"] > 1000, c["cone_background"]
    assert c["cone_rays"] == ((p.width + tile[0] - 1) // tile[0]) * ((p.height + tile[1] - 1) // tile[1])
    # the step counts of the background (the cone's counts where the cone
    # decided, the march's elsewhere) against the oracle's
    steps_d = (cone["gbuf"][bg, 1] >> 16) & 0x3FF
    steps_o = (ref["gbuf"][bg, 1] >> 16) & 0x3FF
    assert (np.abs(steps_d.astype(int) - steps_o.astype(int)) <= 1).mean() >= 0.999


def test_cone_jumps_meet_the_bar(hip):
    """Mode 2 (the march starts at the tile's cone depth): against the
    oracle's march with the same cone grid, the plain bar."""
    sc = H.scene_bulb(320, 180, seed=6)
    p = H.params_from_scene(sc)
    got = hip.march(p, opts=_opts_mode2())
    c = H.counters()
    ref = O.render(O.params(sc), threads=0, cone=(8, 8, 2))
    m = parity.compare(got, ref, p.width, p.height, p.step_width, 2.0)
    assert c["jumps"] > 1000, c
    assert m["mask_agree"] >= 0.999 and m["depth_ok_frac"] >= 0.999, m


def _opts_mode2():
    o = H.options(cone=True)
    o.mlp_cone = 2
    return o


@pytest.mark.parametrize("tile,mode", [((8, 8), 1), ((16, 16), 1), ((8, 8), 2)])
def test_batch_with_cones_equals_single_frames(hip, tile, mode):
    """A batch of uneven frames with the cone rays (k_cone, k_cone_bg's
    background tiles and tile list, the listed march) equals the frames
    marched one by one: 16 x 16 tiles take four wave passes a background
    tile, and mode 2's jumps start from the listed tiles' cone depths."""
    scenes_ = H.animation(5, seed=7, sizes=((96, 72), (160, 120), (200, 150)))
    arr = H.params_array(scenes_)
    n = sum(s.width * s.height for s in scenes_)
    bufs, out = device_out(n)

    def opts():
        o = H.options(cone=True, tile=tile)
        o.mlp_cone = mode
        return o

    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), H.C.byref(opts())), "batch")
    cb = H.counters()
    got = host(bufs)
    off = 0
    bg = 0
    for i, s in enumerate(scenes_):
        m = s.width * s.height
        one = hip.march(arr[i], opts=opts())
        bg += H.counters()["cone_background"]
        assert np.array_equal(got["depth"][off:off + m], one["depth"]), i
        assert np.array_equal(got["gbuf"][off * 5:(off + m) * 5].view(np.uint32), one["gbuf"].ravel()), i
        off += m
    assert cb["cone_background"] == bg > 0


@pytest.mark.parametrize("tile,rnd", [((8, 8), False), ((16, 16), False), ((8, 8), True)])
def test_cone_background_tiles_hand_off_starts_in_the_guard(hip, tile, rnd):
    """With a guard band so wide that every start's DE lies inside it, each
    pixel of a background tile goes from k_cone_bg to the f64 finish at its
    start (as k_march32 hands off a start in the guard), and the picture --
    every output word -- is the plain march's with the same guard; with the
    random first step the hand-off carries the pixel's generator state."""
    sc = H.scene_bulb(256, 136, seed=4)  # (rows of 32 tiles of 8 x 8, taken in fours; 136 rows: ragged 16 x 16)
    sc.first_step_random = 1 if rnd else 0
    p = H.params_from_scene(sc)
    plain = hip.march(p, opts=H.options(tile=tile, guard_band=1e4))
    cone = hip.march(p, opts=H.options(tile=tile, guard_band=1e4, cone=True))
    c = H.counters()
    assert c["cone_rays"] > 0 and c["cone_background"] == 0, c
    for k in ("depth", "depth64", "iters", "normals", "gbuf"):
        assert np.array_equal(cone[k], plain[k]), k
