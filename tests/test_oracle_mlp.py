"""The numpy restatement of the surrogate network (oracle/mlp.py), pinned by
mathematics: its float64 gradient equals central finite differences of its
loss, and its BF16 rounding is round-to-nearest-even on known values."""
import numpy as np

from oracle import mlp


def _small_theta(rng):
    layers = []
    for k, n in zip(mlp.DIMS[:-1], mlp.DIMS[1:]):
        layers.append((rng.normal(0, np.sqrt(2.0 / k), (k, n)), rng.normal(0, 0.1, n)))
    return mlp.pack(layers)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = _small_theta(rng)
    x = rng.normal(size=(8, 14))
    y = rng.normal(size=8)
    loss, g = mlp.loss_and_grad(theta, x, y, bf16_mode=False)
    # probe parameters of every layer, weights and biases
    offs, o = [], 0
    for k, n in zip(mlp.DIMS[:-1], mlp.DIMS[1:]):
        offs += [o, o + k * n - 1, o + k * n]  # first and last weight, first bias
        o += k * n + n
    offs += list(rng.integers(0, len(theta), 20))
    h = 1e-6
    for i in offs:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp.loss_and_grad(tp, x, y, False)[0] - mlp.loss_and_grad(tm, x, y, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_parameter_count():
    assert sum(k * n + n for k, n in zip(mlp.DIMS[:-1], mlp.DIMS[1:])) == 201730


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e38, 1e-40, 0.0], np.float32)
    r = mlp.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0          # halfway between 1 and 1.0078125: to even
    assert r[2] == 1.0078125    # above halfway: up
    assert r[3] == -2.5
    assert np.isfinite(r[4]) and abs(r[4] - 3.0e38) / 3.0e38 < 1e-2
    assert r[6] == 0.0
