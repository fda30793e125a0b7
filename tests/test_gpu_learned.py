"""The learned warm start: a network trained on the plain marcher's (ray,
hit distance) pairs, then the march with and without it on rays it has not
seen (the same view at another resolution of the same aspect: the start
plane's size follows the aspect, so another aspect asks the network for
ray origins it was never shown), both against the f64 oracle.

The bar for the learned march is the mask (>= 99.9 %) and the depth (p99 of
the relative error <= 1e-4 of the scene scale, >= 99 % of the pixels
within it).  A jump changes the march's sample points, so the bisection
ends at another point of the same DE = stop crossing: the depth moves by
the reference's own bisection uncertainty (about a 2^-7 share of the last
step), and near thin features another sample placement can meet another
feature -- the reference's own march is that sensitive to where its samples
fall.  The normal there is not held to the plain march's bar."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(hip):
    train = [H.params_from_scene(H.scene_bulb(480, 270, seed=1))]
    net, curve = H.train_warm_start(train, steps=2000, batch=8192, lr=3e-3, seed=0)
    assert curve[-100:].mean() < 0.1 * curve[:20].mean()
    yield net
    net.close()


def test_the_trained_network_saves_steps_within_the_bar(hip, trained):
    sc = H.scene_bulb(640, 360, seed=1)
    p = hip.params_from_scene(sc)
    plain = hip.march(p)
    learned = hip.march(p, opts=H.options(mlp=trained, conf_miss=0.02))
    cp, cl = plain["counters"], learned["counters"]
    # the gain: most rays start where the network puts them, and the FP32
    # march takes far fewer steps and iterations
    assert cl["jumps"] > 0.9 * cp["pixels"], cl
    assert cl["rollbacks"] < 0.05 * cl["jumps"], cl
    assert cl["steps"] < 0.4 * cp["steps"], (cl["steps"], cp["steps"])
    assert cl["its_p8_f32"] < 0.65 * cp["its_p8_f32"], (cl["its_p8_f32"], cp["its_p8_f32"])
    r = O.render(O.params(sc), threads=0)
    m = parity.compare(learned, r, p.width, p.height, p.step_width, 2.0)
    assert m["mask_agree"] >= 0.999, m
    assert m["depth_rel_p99"] <= 1e-4 and m["depth_ok_frac"] >= 0.99, m
    mp = parity.compare(plain, r, p.width, p.height, p.step_width, 2.0)
    assert mp["depth_ok_frac"] >= 0.999


def test_an_untrained_network_gains_nothing_and_breaks_nothing(hip, trained):
    """The same march with a freshly initialised network: its jumps are
    checked like any other, so the picture keeps the bar, and what the
    trained network saves is the training's."""
    sc = H.scene_bulb(320, 180, seed=1)
    p = hip.params_from_scene(sc)
    raw = H.Mlp(seed=11)
    try:
        plain = hip.march(p)["counters"]
        g_raw = hip.march(p, opts=H.options(mlp=raw, conf_miss=0.02))
        g_tr = hip.march(p, opts=H.options(mlp=trained, conf_miss=0.02))
    finally:
        raw.close()
    r = O.render(O.params(sc), threads=0)
    m = parity.compare(g_raw, r, p.width, p.height, p.step_width, 2.0)
    assert m["mask_agree"] >= 0.995, m
    saved_raw = 1.0 - g_raw["counters"]["steps"] / plain["steps"]
    saved_tr = 1.0 - g_tr["counters"]["steps"] / plain["steps"]
    assert saved_tr > 0.5 and saved_tr > saved_raw + 0.3, (saved_raw, saved_tr)


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
