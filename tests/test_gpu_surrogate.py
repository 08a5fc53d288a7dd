"""The surrogate network on the GPU: forward pass and gradients against the
numpy oracle (oracle/mlp.py, itself checked against finite differences),
training on the engine's own ray-shooting output, what the trained network's
uncertainty gate gains over the methods without it, and the gated
surrogate + ray-shooting light curve against plain ray shooting."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from triplelens_amd import surrogate
    return surrogate


def _rand_xy(rng, n):
    X = rng.normal(size=(n, 14)).astype(np.float32)
    y = (0.3 * np.tanh(X[:, 0]) - 0.2 * X[:, 9] + 0.05 * rng.normal(size=n)).astype(np.float32)
    return X, y


def test_forward_matches_oracle(S):
    rng = np.random.default_rng(1)
    net = S.Net(seed=5)
    theta = net.params()
    assert theta.size == 201730
    X, _ = _rand_xy(rng, 300)
    got = net.predict_raw(X)
    want, _ = mlp.forward(theta, X, bf16_mode=True)
    assert np.max(np.abs(got - want)) <= 2e-2 * (1 + np.max(np.abs(want)))
    assert np.corrcoef(got.ravel(), want.ravel())[0, 1] > 0.9999


def test_gradients_match_oracle(S):
    rng = np.random.default_rng(2)
    net = S.Net(seed=9)
    theta = rng.normal(0, 0.05, S.Net.param_count()).astype(np.float32)
    net.set_params(theta)
    X, y = _rand_xy(rng, 256)
    loss, g = net.grads(X, y)
    loss_o, g_o = mlp.loss_and_grad(net.params(), X, y, bf16_mode=True)
    assert abs(loss - loss_o) <= 1e-3 * abs(loss_o)
    cos = float(g @ g_o / (np.linalg.norm(g) * np.linalg.norm(g_o)))
    rel = float(np.linalg.norm(g - g_o) / np.linalg.norm(g_o))
    assert cos > 0.999 and rel < 5e-2, (cos, rel)
    # every layer's weight and bias gradient, not only the dominant ones
    o = 0
    for k, n in zip(mlp.DIMS[:-1], mlp.DIMS[1:]):
        for sl in (slice(o, o + k * n), slice(o + k * n, o + k * n + n)):
            a, b = g[sl], g_o[sl]
            assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b) + 1e-7
        o += k * n + n


def test_training_reduces_loss(S):
    rng = np.random.default_rng(3)
    net = S.Net(seed=4)
    X, y = _rand_xy(rng, 8192)
    net.fit_scaling(X, y)
    losses = net.train(X, y, steps=600, batch=256, lr=2e-3, seed=7)
    assert np.mean(losses[-50:]) < np.mean(losses[:50]) - 1.0
    pred = net.predict(X)
    assert np.sqrt(np.mean((pred[:, 0] - y) ** 2)) < 0.1


def _curves(seeds, n_points, ppr):
    from triplelens_amd import engine as E
    from triplelens_amd import workloads as W

    def ca_fn(a, b, th, m2, m3):
        return E.LightCurve(a, b, th, m2, m3, 5e-4, 1, ppr).cc_ca()[1]

    return [W.c2(seed=s, n_points=n_points, ppr=ppr, ca_fn=ca_fn) for s in seeds]


def _zoom(S, seeds):
    """C2-distribution curves cut to 100 source radii about the central
    caustic crossing, where the finite-source corrections are."""
    return [S.caustic_zoom(c, 0.025) for c in _curves(seeds, 512, 100)]


@pytest.fixture(scope="module")
def trained(S):
    """A network trained on the engine's ray shooting of a few curves
    (near-caustic and full-length C2 curves at pixel rho/100), its gate
    calibrated on other curves: the training pipeline at test size."""
    X, y, raw = S.training_set(_zoom(S, range(1000, 1048)) + _curves(range(4000, 4016), 512, 100))
    Xc, yc, rc = S.training_set(_zoom(S, range(3000, 3016)) + _curves(range(5000, 5008), 512, 100))
    sur = S.Surrogate(S.Net(seed=1), gate_radii=2.0)
    losses = sur.fit(X, y, raw["cross"], steps=4000, batch=512, lr=1e-3, seed=2)
    sigma_max = sur.calibrate(Xc, yc, rc["cross"], bound=1e-3)
    return sur, losses, sigma_max, (Xc, yc, rc)


@pytest.fixture(scope="module")
def shipped(S):
    """The C2 model shipped with the package (python -m
    triplelens_amd.train_surrogate: trained and calibrated on ray-shot C2
    geometries with seeds >= 10000 at pixel rho/200)."""
    import os
    path = os.path.join(os.path.dirname(S.__file__), "data", "surrogate_c2.npz")
    return S.Surrogate.load(path)


def _zoom200(S, seeds):
    return [S.caustic_zoom(c, 0.025) for c in _curves(seeds, 512, 200)]


def test_training_pipeline_on_engine_output(S, trained):
    """Training on ray-shot data lowers the loss, and the calibrated gate
    keeps every accepted calibration position within the bound / safety."""
    sur, losses, sigma_max, (Xc, yc, rc) = trained
    assert np.isfinite(losses).all() and losses[-200:].mean() < losses[:200].mean()
    assert np.all(sigma_max > 0)
    mu, sigma = sur.predict(Xc)
    acc = sur.accepts(Xc, rc["cross"], sigma)
    assert acc.mean() > 0.3
    assert np.abs(np.expm1(mu - yc))[acc].max() <= 0.5e-3


def test_network_gate_gains_over_the_methods_without_it(S, shipped):
    """On held-out near-caustic C2 curves (pixel rho/200), at <= 1e-3
    relative error against ray shooting, the shipped network's gate accepts
    at least as many positions as the methods without the network can, and
    far more than A_point alone: the methods without it are the
    hexadecapole estimate behind the same physical gates (no disk sample
    across a caustic, converged expansion) and A_point alone, each with the
    tightest distance-to-caustic gate that keeps it within the bound on
    these very curves (a choice made after the fact, in their favour).  The
    printed line gives the three acceptance rates."""
    sur = shipped
    X, y, raw = S.training_set(_zoom200(S, range(200, 216)))
    mu, sigma = sur.predict(X)
    d = np.abs(raw["d"])
    err_net = np.abs(np.expm1(mu - y))
    acc_net = sur.accepts(X, raw["cross"], sigma)
    err_hex = np.abs(np.expm1(X[:, S.F_HEX] - y))
    err_point = np.abs(np.expm1(-y))
    physical = sur.candidates(X, raw["cross"]) | (d < sur.gate_radii)

    def distance_gate(err, extra):
        for g in (2, 3, 5, 7, 10, 15, 20, 30, 50, 100):
            sel = (d >= g) & extra
            if err[sel].max(initial=0.0) <= 1e-3:
                return float(np.mean(sel)), g
        return 0.0, None

    acc_hex, g_hex = distance_gate(err_hex, physical)
    acc_point, g_point = distance_gate(err_point, np.ones_like(physical))
    print("held-out %d positions: network gate accepts %.3f (max err %.2e); hexadecapole "
          "behind the physical gates %.3f (|d| >= %s rho); A_point alone %.3f (|d| >= %s rho)"
          % (len(y), acc_net.mean(), err_net[acc_net].max(initial=0.0), acc_hex, g_hex,
             acc_point, g_point))
    assert err_net[acc_net].max(initial=0.0) <= 1e-3
    assert acc_net.mean() >= acc_hex
    assert acc_net.mean() >= 2.0 * acc_point


def test_surrogate_light_curve_against_ray_shooting(S, shipped):
    """Gated surrogate + ray shooting on held-out C2 curves: fallback
    positions are ray-shot from the union of four split seed searches, so
    they count at least what plain ray shooting with the first split search
    counts (equal at most positions); against plain ray shooting with the
    reference's single root chain every position stays within the 1e-3
    relative-error bound."""
    from triplelens_amd import engine as E
    sur = shipped
    sur.audit = 0.05
    for c in _zoom200(S, range(300, 302)) + _curves(range(310, 312), 1024, 200):
        lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        lc.set_limb_darkening("linear", c.u)
        A, info = sur.light_curve(lc, c, np.random.default_rng(1))
        split = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        split.set_limb_darkening("linear", c.u)
        split.set_seed_split(sur.seed_parts(c.n_points))
        A_split, _, _ = split.lc_irs(*c.traj())
        fb = info["fallback"]
        assert np.all(A[fb] >= A_split[fb])  # a superset of the seeds
        assert np.mean(A[fb] == A_split[fb]) > 0.9
        ref = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        ref.set_limb_darkening("linear", c.u)
        A_ref, _, _ = ref.lc_irs(*c.traj())
        rel = np.abs(A - A_ref) / A_ref
        print("fallback %.1f%%, max rel err of accepted %.2e, audit max %.2e"
              % (100 * fb.mean(), rel[~fb].max() if (~fb).any() else 0.0,
                 info["audit_rel_err"].max() if info["audit_rel_err"].size else 0.0))
        assert rel.max() <= 1e-3


def test_batched_light_curves_equal_single_ones(S, shipped):
    """light_curves over several curves (one seed launch, one network call,
    one fill) gives each curve what light_curve gives it alone."""
    from triplelens_amd import engine as E
    sur = shipped
    sur.audit = 0.05
    curves = _zoom200(S, range(320, 322)) + _curves(range(330, 331), 512, 200)

    def new(c):
        lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        lc.set_limb_darkening("linear", c.u)
        return lc

    rng = np.random.default_rng(4)
    single = [sur.light_curve(new(c), c, rng) for c in curves]
    batched = sur.light_curves([new(c) for c in curves], curves, np.random.default_rng(4))
    for (A1, i1), (A2, i2) in zip(single, batched):
        assert np.array_equal(i1["fallback"], i2["fallback"])
        assert np.array_equal(i1["audited"], i2["audited"])
        assert np.array_equal(A1, A2)
