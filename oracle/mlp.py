"""CPU restatement of the surrogate network (triplelens_amd/csrc/tl_mlp.hip,
tl_surrogate.cpp) in numpy.

TEST INFRASTRUCTURE ONLY: the checker for the GPU network's forward pass and
gradients; the product never imports it.

Network: 14 -> 256 -> 256 -> 256 -> 256 -> 2, GELU, Gaussian NLL loss
0.5 ((y - mu) e^-s)^2 + s (s clamped to [-12, 12]), mean over the batch.

Two modes:
  * bf16=True reproduces the GPU's rounding points: BF16 inputs, weights,
    activations and backward signals, FP32 products and sums;
  * bf16=False is plain float64, the mathematics itself (its gradient is
    checked against finite differences in tests/test_oracle_mlp.py).

Parameters use the logical order of tl_mlp_get_params: W1 (14 x 256), b1,
W2, b2, W3, b3, W4, b4, W5 (256 x 2), b5.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

DIMS = [14, 256, 256, 256, 256, 2]


def unpack(theta):
    """Flat logical vector -> [(W, b)] per layer."""
    out, o = [], 0
    for k, n in zip(DIMS[:-1], DIMS[1:]):
        W = theta[o:o + k * n].reshape(k, n)
        o += k * n
        b = theta[o:o + n]
        o += n
        out.append((W, b))
    assert o == len(theta)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = r.astype(np.uint32)
    nonfinite = (u & 0x7F800000) == 0x7F800000
    r[nonfinite] = (u[nonfinite] & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(theta, x, bf16_mode=True):
    """Returns (out (n, 2), cache)."""
    dt = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda v: v)
    layers = unpack(np.asarray(theta, dt))
    h = rnd(np.asarray(x, dt))
    hs, zs = [h], []
    for li, (W, b) in enumerate(layers):
        z = (h @ rnd(W)).astype(dt) + b
        if li < len(layers) - 1:
            zs.append(z)
            h = rnd(gelu(z).astype(dt))
            hs.append(h)
        else:
            out = z
    return out, (layers, hs, zs)


def loss_and_grad(theta, x, y, bf16_mode=True):
    """Mean Gaussian NLL and its gradient (logical order)."""
    dt = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda v: v)
    out, (layers, hs, zs) = forward(theta, x, bf16_mode)
    y = np.asarray(y, dt)
    n = len(y)
    mu, s0 = out[:, 0], out[:, 1]
    s = np.clip(s0, -12.0, 12.0)
    r = y - mu
    e = np.exp(-2.0 * s)
    loss = float(np.mean(0.5 * r * r * e + s))
    dz = np.zeros_like(out)
    dz[:, 0] = -r * e / n
    dz[:, 1] = np.where(s0 == s, (1.0 - r * r * e) / n, 0.0)
    dz = rnd(dz)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        gW = (hs[li].T @ dz).astype(dt)
        gb = dz.sum(axis=0).astype(dt)
        grads[li] = (gW, gb)
        if li == 0:
            break
        dh = (dz @ rnd(W).T).astype(dt)
        dz = rnd((dh * gelu_d(zs[li - 1])).astype(dt))
    return loss, pack(grads)
