"""The warm-start network on the GPU: its FP32 forward and gradient against
the f64 CPU reference (oracle/mlp_ref.py), the BF16 MFMA inference against
the FP32 forward, and training that lowers the loss."""
import numpy as np
import pytest

import mb3d_hip as H
import mlp_ref as R

pytestmark = pytest.mark.gpu


def rays_targets(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * rng.uniform(1.5, 3.5, (n, 1))
    d = -o / np.linalg.norm(o, axis=1, keepdims=True) + 0.4 * rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a unit sphere: the exact hit distance where the ray meets it
    b = (o * d).sum(1)
    disc = b * b - ((o * o).sum(1) - 1.0)
    hit = disc > 0
    t = np.where(hit, -b - np.sqrt(np.maximum(disc, 0)), 0.0)
    tg = np.stack([t / 4.0, hit.astype(np.float64)], 1)
    return np.concatenate([o, d], 1), tg


def test_forward_f32_matches_the_reference(hip):
    net = H.Mlp(seed=1)
    theta = R.init(3)
    net.params(theta.astype(np.float32))
    rays, _ = rays_targets(1000, 1)
    got = net.infer(rays, mfma=False)
    want, _ = R.forward(theta.astype(np.float32).astype(np.float64), R.encode(rays.astype(np.float32)))
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-5 * scale, np.abs(got - want).max()
    net.close()


def test_gradient_matches_the_reference(hip):
    net = H.Mlp(seed=1)
    theta = R.init(4).astype(np.float32)
    rng = np.random.default_rng(0)
    theta[-2:] = rng.normal(size=2).astype(np.float32)  # biases off zero
    net.params(theta)
    rays, tg = rays_targets(700, 2)  # several row chunks and a ragged last one
    loss, g = net.grad(rays, tg)
    lw, gw = R.gradient(theta.astype(np.float64), R.encode(rays.astype(np.float32)), tg.astype(np.float32))
    assert abs(loss - lw) <= 1e-5 * max(1.0, abs(lw)), (loss, lw)
    o = 0
    for s in R.SHAPES:
        n = int(np.prod(s))
        a, b = g[o:o + n], gw[o:o + n]
        err = np.abs(a - b).max()
        assert err <= 1e-4 * max(np.abs(b).max(), 1e-6) + 1e-7, (s, o, err, np.abs(b).max())
        o += n
    # the padding rows of W1 get gradients but stay zero in the weights
    assert (net.params()[78 * 64:80 * 64] == 0).all()
    net.close()


def test_gradient_is_deterministic(hip):
    net = H.Mlp(seed=5)
    rays, tg = rays_targets(3000, 3)
    a = net.grad(rays, tg)
    b = net.grad(rays, tg)
    # the gradient sums its chunks in order; the reported loss adds one
    # partial per wave atomically, so only its rounding may differ
    assert np.array_equal(a[1], b[1])
    assert abs(a[0] - b[0]) <= 1e-6 * abs(a[0])
    net.close()


def test_mfma_inference_matches_f32(hip):
    net = H.Mlp(seed=2)
    rays, tg = rays_targets(20000, 4)
    net.train(rays, tg, steps=300, batch=2048, lr=3e-3)
    for n in (64, 1000, 4099):  # whole waves, a ragged last wave
        r = rays[:n]
        f32 = net.infer(r, mfma=False)
        bf = net.infer(r, mfma=True)
        for c in range(2):
            scale = np.abs(f32[:, c]).max()
            err = np.abs(bf[:, c] - f32[:, c])
            # BF16 operands (8-bit mantissa), FP32 accumulation over <= 80 terms
            assert err.max() <= 0.03 * scale, (n, c, err.max(), scale)
            assert np.corrcoef(bf[:, c], f32[:, c])[0, 1] > 0.999
    net.close()


def test_training_lowers_the_loss(hip):
    net = H.Mlp(seed=3)
    rays, tg = rays_targets(40000, 5)
    test_r, test_t = rays_targets(5000, 6)
    h = test_t[:, 1] > 0.5

    def dist_err():
        return np.median(np.abs(net.infer(test_r, mfma=True)[h, 0] - test_t[h, 0]) * 4.0)

    before = dist_err()
    curve = net.train(rays, tg, steps=1500, batch=4096, lr=3e-3)
    assert np.isfinite(curve).all()
    assert curve[-50:].mean() < 0.25 * curve[:10].mean(), (curve[:10].mean(), curve[-50:].mean())
    # held-out rays: the hit classification mostly right, the distance closer
    out = net.infer(test_r, mfma=True)
    acc = ((out[:, 1] > 0) == h).mean()
    assert acc > 0.93, acc
    after = dist_err()
    assert after < 0.5 * before and after < 0.25, (before, after)
    net.close()


def test_params_roundtrip_and_bad_count(hip):
    net = H.Mlp(seed=7)
    p = net.params()
    assert p.shape == (H.MLP_PARAMS,) and np.isfinite(p).all()
    q = (p * 0.5).astype(np.float32)
    assert np.array_equal(net.params(q), q)
    lib = H.load()
    assert lib.mb3d_hip_mlp_params(net.h, None, None, 5) < 0
    net.close()
