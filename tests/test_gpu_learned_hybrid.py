"""The learned warm start on the hybrid chain (Mandelbulb p8 x1 + Amazing
Box scale 2 x2, numeric DE) at normal zoom: a network trained on ~1M (ray,
hit distance) pairs that the plain marcher gives for 12 seeded cameras,
used on a camera it was not trained on, with the cone rays as the verified
miss decision.  Both marches run the FP32 iterate (the automatic precision
would march this numeric DE in FP64, where the warm start does not apply).

Held: the learned march keeps the plain FP32 march's own mask and depth
error against the f64 oracle, and takes fewer FP32 march steps."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu


def test_hybrid_hit_jumps_keep_the_plain_fp32_bar(hip):
    f32 = H.options(precision=H.PREC_F32)
    train = [H.params_from_scene(H.scene_hybrid(384, 216, seed=s)) for s in range(10, 22)]
    rs, ts = zip(*(H.training_pairs(p, opts=f32) for p in train))
    rays, tg = np.concatenate(rs), np.concatenate(ts)
    assert len(rays) > 900_000
    perm = np.random.default_rng(0).permutation(len(rays))
    net = H.Mlp(0)
    try:
        curve = net.train(rays[perm], tg[perm], steps=3000, batch=8192, lr=3e-3)
        assert curve[-100:].mean() < 0.1 * curve[:20].mean()
        sc = H.scene_hybrid(640, 360, seed=1)
        p = hip.params_from_scene(sc)
        plain = hip.march(p, opts=f32)
        learned = hip.march(p, opts=H.options(precision=H.PREC_F32, mlp=net, cone=True))
        r = O.render(O.params(sc), threads=0)
        mp = parity.compare(plain, r, p.width, p.height, p.step_width, 2.0)
        ml = parity.compare(learned, r, p.width, p.height, p.step_width, 2.0)
        cp, cl = plain["counters"], learned["counters"]
        print(f"plain f32: {cp['kernel_ms']:.3f} ms, {cp['steps']} steps, mask {mp['mask_agree']:.5f}, "
              f"depth ok {mp['depth_ok_frac']:.5f}; learned (hit jumps): {cl['kernel_ms']:.3f} ms, "
              f"{cl['steps']} steps, {cl['jumps']} jumps, {cl['rollbacks']} rollbacks, "
              f"mask {ml['mask_agree']:.5f}, depth ok {ml['depth_ok_frac']:.5f}")
        assert cl["jumps"] > 0
        assert ml["mask_agree"] >= min(0.999, mp["mask_agree"] - 0.0005), (ml, mp)
        assert ml["depth_ok_frac"] >= mp["depth_ok_frac"] - 0.01, (ml, mp)
        # the saving on the unseen camera
        assert cl["steps"] < 0.8 * cp["steps"], (cl["steps"], cp["steps"])
    finally:
        net.close()
