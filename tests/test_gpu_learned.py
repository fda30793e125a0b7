"""The learned warm start: a network trained on the plain marcher's (ray,
hit distance) pairs of 48 seeded cameras (240 x 135 each, about 1.5 M
pairs), then the march with and without it on cameras it was not trained
on, both against the f64 oracle.

The network sees a ray as its point of closest approach to the origin and
its direction (a camera-independent form), and proposes the hit's distance;
a ray it calls a hit jumps there less a margin, and the exact DE at the
landing keeps or rolls back the jump.  The bar for the learned march is the
target's: mask >= 99.9 %, p99 of the relative depth error <= 1e-4 of the
scene scale.  A jump changes the march's sample points, so the bisection
ends elsewhere on the same DE = stop crossing, and near thin features
another sample placement can meet another feature: the depth of about 0.7 %
of the pixels moves by more than 1e-4, and the normals, which turn within a
pixel on a fractal, are not held to the plain march's bar."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu


TRAIN_SEEDS = range(10001, 10049)


@pytest.fixture(scope="module")
def trained(hip):
    train = [H.params_from_scene(H.scene_bulb(240, 135, seed=s)) for s in TRAIN_SEEDS]
    net, curve = H.train_warm_start(train, steps=2000, batch=8192, lr=3e-3, seed=0)
    assert curve[-100:].mean() < 0.05 * curve[:20].mean()
    yield net
    net.close()


@pytest.mark.parametrize("seed", [1, 2])
def test_the_trained_network_saves_steps_on_unseen_cameras(hip, trained, seed):
    sc = H.scene_bulb(640, 360, seed=seed)
    p = hip.params_from_scene(sc)
    plain = hip.march(p)
    learned = hip.march(p, opts=H.options(mlp=trained))
    cp, cl = plain["counters"], learned["counters"]
    hits = cp["hits"]
    # most surface pixels jump, few jumps are rolled back, and the FP32
    # march takes fewer steps
    assert cl["jumps"] > 0.8 * hits and cl["rollbacks"] < 0.05 * cl["jumps"], (cl, hits)
    assert cl["steps"] < cp["steps"] - 5 * hits, (cl["steps"], cp["steps"], hits)
    r = O.render(O.params(sc), threads=0)
    m = parity.compare(learned, r, p.width, p.height, p.step_width, 2.0)
    assert m["mask_agree"] >= 0.999, m
    assert m["depth_rel_p99"] <= 1e-4 and m["depth_ok_frac"] >= 0.98, m
    mp = parity.compare(plain, r, p.width, p.height, p.step_width, 2.0)
    assert mp["depth_ok_frac"] >= 0.999


def test_an_untrained_network_gains_nothing_and_breaks_nothing(hip, trained):
    """The same march with a freshly initialised network: its jumps are
    checked like any other, so the picture keeps the mask, and what the
    trained network saves is the training's."""
    sc = H.scene_bulb(320, 180, seed=3)
    p = hip.params_from_scene(sc)
    raw = H.Mlp(seed=11)
    try:
        plain = hip.march(p)["counters"]
        g_raw = hip.march(p, opts=H.options(mlp=raw))
        g_tr = hip.march(p, opts=H.options(mlp=trained))
    finally:
        raw.close()
    r = O.render(O.params(sc), threads=0)
    m = parity.compare(g_raw, r, p.width, p.height, p.step_width, 2.0)
    assert m["mask_agree"] >= 0.999, m
    saved_raw = 1.0 - g_raw["counters"]["steps"] / plain["steps"]
    saved_tr = 1.0 - g_tr["counters"]["steps"] / plain["steps"]
    assert saved_tr > saved_raw + 0.05, (saved_raw, saved_tr)


def test_the_cone_keeps_the_learned_picture(hip, trained):
    """With the cone rays the tiles they find empty are background without
    the network (its wave is skipped): the learned picture, and fewer
    steps."""
    sc = H.scene_bulb(640, 360, seed=4)
    p = hip.params_from_scene(sc)
    a = hip.march(p, opts=H.options(mlp=trained))
    b = hip.march(p, opts=H.options(mlp=trained, cone=True))
    assert b["counters"]["cone_background"] > 0.3 * p.width * p.height
    assert (a["depth"] == b["depth"]).mean() >= 0.9999
    assert b["counters"]["steps"] < 0.7 * a["counters"]["steps"]


def test_jumps_off_is_the_plain_march(hip, trained):
    """With both jump kinds turned off the learned kernel marches exactly
    as the plain one."""
    sc = H.scene_bulb(320, 240, seed=2)
    p = hip.params_from_scene(sc)
    plain = hip.march(p)
    off = hip.march(p, opts=H.options(mlp=trained, conf_hit=2.0))
    assert off["counters"]["jumps"] == 0
    for k in ("depth", "iters", "gbuf"):
        assert np.array_equal(off[k], plain[k]), k


def test_learned_batch_equals_learned_single_frames(hip, trained):
    import torch

    scenes = H.animation(4, seed=3, sizes=((96, 72), (160, 120), (64, 48)))
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    dev = torch.device("cuda:0")
    depth = torch.zeros(n, dtype=torch.float32, device=dev)
    iters = torch.zeros(n, dtype=torch.int16, device=dev)
    out = H.Out(depth.data_ptr(), 0, iters.data_ptr(), 0, 0, 1, 0)
    opts = H.options(mlp=trained)
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), H.C.byref(opts)), "batch")
    assert H.counters()["jumps"] > 0
    torch.cuda.synchronize()
    got = depth.cpu().numpy()
    off = 0
    for i, s in enumerate(scenes):
        m = s.width * s.height
        one = hip.march(arr[i], opts=opts, outputs=("depth",))
        assert np.array_equal(got[off:off + m], one["depth"]), i
        off += m
