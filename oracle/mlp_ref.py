"""TEST INFRASTRUCTURE: the f64 CPU reference of the warm-start MLP.

The network has no counterpart in the reference (it is the learned step
proposer of this backend).  This module is its checker: the sin/cos ray
encoding, the forward pass, the loss and its analytic gradient in float64,
the gradient pinned by central finite differences in the tests.

Layout (the same as the HIP kernels'): inputs x = (origin / 4, unit
direction); features: for each input dim i and frequency k = 0..5,
sin(2^k pi x_i) and cos(2^k pi x_i) (72), then the 6 raw inputs, padded with 2
zeros to 80; layers 80 -> 64 -> 64 -> 64 -> 2 with ReLU between; outputs
(t / 4, hit logit).  Parameters in one flat float vector:
W1 (80 x 64, row-major: input-major), b1, W2 (64 x 64), b2, W3, b3,
W4 (64 x 2), b4.
Loss per ray: BCE(logit, hit) + hit * a * (t_hat - t)^2 with a = 4 where
t_hat > t (an overshoot) and 1 below; averaged over the batch.
"""
from __future__ import annotations

import numpy as np

N_IN, N_FREQ, K1, H, N_OUT = 6, 6, 80, 64, 2
SHAPES = [(K1, H), (H,), (H, H), (H,), (H, H), (H,), (H, N_OUT), (N_OUT,)]
N_PARAMS = int(sum(np.prod(s) for s in SHAPES))
OVER = 4.0


def unpack(theta):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(theta[o:o + n], np.float64).reshape(s))
        o += n
    return out


def encode(rays):
    """rays: n x 6 (origin xyz, unit direction xyz) -> n x 80 features."""
    r = np.asarray(rays, np.float64)
    x = np.concatenate([r[:, :3] / 4.0, r[:, 3:6]], axis=1)
    f = np.zeros((len(r), K1))
    c = 0
    for i in range(N_IN):
        for k in range(N_FREQ):
            a = np.pi * (2.0 ** k) * x[:, i]
            f[:, c] = np.sin(a)
            f[:, c + 1] = np.cos(a)
            c += 2
    f[:, c:c + N_IN] = x
    return f


def forward(theta, feats):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(theta)
    z1 = feats @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ W3 + b3
    h3 = np.maximum(z3, 0.0)
    out = h3 @ W4 + b4
    return out, (feats, z1, h1, z2, h2, z3, h3)


def loss_and_dout(out, targets):
    """targets: n x 2 (t / 4, hit 0/1)."""
    t, hit = targets[:, 0], targets[:, 1]
    th, lg = out[:, 0], out[:, 1]
    n = len(out)
    p = 1.0 / (1.0 + np.exp(-lg))
    bce = np.maximum(lg, 0) - lg * hit + np.log1p(np.exp(-np.abs(lg)))
    e = th - t
    a = np.where(e > 0, OVER, 1.0)
    loss = (bce + hit * a * e * e).mean()
    d = np.zeros_like(out)
    d[:, 0] = hit * a * 2.0 * e / n
    d[:, 1] = (p - hit) / n
    return loss, d


def gradient(theta, feats, targets):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(theta)
    out, (x, z1, h1, z2, h2, z3, h3) = forward(theta, feats)
    loss, dout = loss_and_dout(out, targets)
    gW4 = h3.T @ dout
    gb4 = dout.sum(0)
    d3 = (dout @ W4.T) * (z3 > 0)
    gW3 = h2.T @ d3
    gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * (z2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g


def loss_only(theta, feats, targets):
    out, _ = forward(theta, feats)
    return loss_and_dout(out, targets)[0]


def init(seed=0):
    """He-initialised parameters (float64)."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            fan_in = s[0]
            parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), size=s).ravel())
        else:
            parts.append(np.zeros(s))
    theta = np.concatenate(parts)
    # the two padding rows of W1 stay zero
    W1 = theta[:K1 * H].reshape(K1, H)
    W1[N_IN * N_FREQ * 2 + N_IN:, :] = 0.0
    return theta
