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


def test_the_cone_gate_keeps_the_plain_picture(hip, trained):
    """The network as the cone pass's gate, hit jumps off: a tile whose
    middle ray the network calls a hit marches no cone ray, its rays march
    plainly -- the plain depths, iterations and normals bit for bit, with
    fewer cone rays (the network's verified gain: a wrong call costs time,
    never a pixel).  The G-buffer records are the plain march's or, for the
    background of a tile whose cone ray went on, the cone's (its dynamic
    fog count, as the reference's cone march leaves it)."""
    sc = H.scene_bulb(640, 360, seed=5)
    p = hip.params_from_scene(sc)
    plain = hip.march(p)
    cone = hip.march(p, opts=H.options(cone=True))
    gated = hip.march(p, opts=H.options(mlp=trained, cone=True, cone_gate=0.9, conf_hit=2.0))
    cc, cg = cone["counters"], gated["counters"]
    for k in ("depth", "depth64", "iters", "normals"):
        assert np.array_equal(gated[k], plain[k]), k
    same = (gated["gbuf"] == plain["gbuf"]).all(1) | (gated["gbuf"] == cone["gbuf"]).all(1)
    assert same.all()
    tiles = (p.width // 8) * (p.height // 8)
    assert cc["cone_rays"] == tiles and cg["mlp_rays"] >= tiles
    assert cg["jumps"] == 0
    # most tiles that hold a surface skip their cone ray; the empty ones
    # keep theirs
    assert cg["cone_rays"] < cc["cone_rays"] - 0.5 * (tiles - cc["cone_background"] / 64), (cc, cg)
    assert cg["cone_background"] >= 0.97 * cc["cone_background"], (cc, cg)


def test_the_cone_gate_in_a_batch_equals_the_plain_batch(hip, trained):
    from test_gpu_batch import device_out, host

    scenes = H.animation(4, seed=7, sizes=((160, 120), (96, 72), (200, 150)))
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    b0, o0 = device_out(n)
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(o0), None), "batch")
    want = host(b0)
    b1, o1 = device_out(n)
    opts = H.options(mlp=trained, cone=True, cone_gate=0.9, conf_hit=2.0)
    plan = H.Plan(arr, o1, opts)
    try:
        for _ in range(2):
            plan.launch(1)
            got = host(b1)
            for k in ("depth", "depth64", "iters", "normals"):
                assert np.array_equal(got[k], want[k]), k
    finally:
        plan.close()


def test_the_cone_gate_needs_the_network_and_the_cone(hip, trained):
    sc = H.scene_bulb(64, 48, seed=1)
    p = hip.params_from_scene(sc)
    for o in (H.options(cone=True, cone_gate=0.9), H.options(mlp=trained, cone_gate=0.9),
              H.options(mlp=trained, cone=True, cone_gate=1.5)):
        with pytest.raises(RuntimeError):
            hip.march(p, opts=o)
