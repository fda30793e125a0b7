"""The f64 MLP reference (the checker of the warm-start network's HIP
kernels) pinned by central finite differences and closed forms."""
import numpy as np

import mlp_ref as M


def batch(n=64, seed=1):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3)) * 2.0
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = rng.uniform(0.2, 1.0, n)
    hit = (rng.uniform(size=n) < 0.5).astype(np.float64)
    return np.concatenate([o, d], 1), np.stack([t, hit], 1)


def test_encoding_closed_form():
    rays = np.array([[4.0, 0.0, -2.0, 0.0, 1.0, 0.0]])
    f = M.encode(rays)[0]
    assert f.shape == (80,)
    # input 0 is x = 1: sin(2^k pi) = 0, cos(2^k pi) = (-1)^(k == 0) ...
    assert np.allclose(f[0:12:2], 0.0, atol=1e-12)
    assert np.allclose(f[1:12:2], [-1, 1, 1, 1, 1, 1])
    assert np.allclose(f[72:78], [1.0, 0.0, -0.5, 0.0, 1.0, 0.0])
    assert (f[78:] == 0).all()


def test_gradient_matches_finite_differences():
    rays, tg = batch()
    feats = M.encode(rays)
    theta = M.init(3)
    # move the biases off zero so every ReLU is away from its kink
    theta += np.random.default_rng(4).normal(0, 0.05, theta.shape)
    _, g = M.gradient(theta, feats, tg)
    rng = np.random.default_rng(5)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        fd = (M.loss_only(theta + e, feats, tg) - M.loss_only(theta - e, feats, tg)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_loss_closed_form_for_a_linear_net():
    """With every weight zero but the output bias, the loss is the BCE of the
    bias and the asymmetric square of bias - t."""
    rays, tg = batch(8)
    theta = np.zeros(M.N_PARAMS)
    theta[-2:] = [0.7, -0.3]
    out, _ = M.forward(theta, M.encode(rays))
    assert np.allclose(out, [[0.7, -0.3]] * 8)
    loss, _ = M.loss_and_dout(out, tg)
    e = 0.7 - tg[:, 0]
    a = np.where(e > 0, 4.0, 1.0)
    p = 1 / (1 + np.exp(0.3))
    bce = -(tg[:, 1] * np.log(p) + (1 - tg[:, 1]) * np.log(1 - p))
    assert abs(loss - (bce + tg[:, 1] * a * e * e).mean()) < 1e-12
