"""The surrogate's physics features on the CPU: the disk moments of the
limb-darkening profile against quadrature, the hexadecapole assembly against
the exact disk mean of fields it must reproduce, and the oracle's disk-sample
magnifications (tlo_ring_amp) against images found independently with numpy's
polynomial roots."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate

from triplelens_amd import surrogate as S


def _profile_mean(f, u):
    """Mean of f(r, t) over the unit disk weighted by 1 - u (1 - sqrt(1 - r^2))."""
    w = lambda r: (1.0 - u * (1.0 - math.sqrt(max(0.0, 1.0 - r * r)))) * r
    num = integrate.dblquad(lambda t, r: f(r, t) * w(r), 0.0, 1.0, 0.0, 2 * math.pi,
                            epsabs=1e-13, epsrel=1e-12)[0]
    den = integrate.quad(lambda r: 2 * math.pi * w(r), 0.0, 1.0, epsabs=1e-14)[0]
    return num / den


@pytest.mark.parametrize("u", [0.0, 0.3, 0.5, 0.8])
def test_disk_moments_match_quadrature(u):
    x2, x4 = S.disk_moments(u)
    assert x2 == pytest.approx(_profile_mean(lambda r, t: r * r, u), rel=1e-10)
    assert x4 == pytest.approx(_profile_mean(lambda r, t: r ** 4, u), rel=1e-10)
    if u == 0.0:
        assert (x2, x4) == (0.5, pytest.approx(1.0 / 3.0))


@pytest.mark.parametrize("u", [0.0, 0.5])
def test_hexadecapole_exact_for_fields_up_to_fourth_order(u):
    """For A = c0 + c2 r^2 + c4 r^4 plus angular modes m = 1, 2, 3 (which
    the four-point means cancel), the expansion is the exact disk mean; a
    dropped term, a wrong weight or a wrong sample radius breaks this."""
    rho = 0.37
    c0, c2, c4, d1, d2, d3 = 2.0, 0.8, -3.1, 0.6, 1.7, -0.9

    def A(r, t):  # r in units of rho
        rr = r * rho
        return (c0 + c2 * rr ** 2 + c4 * rr ** 4 + d1 * rr * math.cos(t)
                + d2 * rr ** 2 * math.cos(2 * t) + d3 * rr ** 3 * math.sin(3 * t))

    c = math.sqrt(0.5)
    off = [(0.5, 0), (0, 0.5), (-0.5, 0), (0, -0.5), (1, 0), (0, 1), (-1, 0), (0, -1),
           (c, c), (-c, c), (-c, -c), (c, -c)]
    ring = np.array([[A(math.hypot(x, y), math.atan2(y, x)) for x, y in off]])
    A_hex, t4 = S.hexadecapole(np.array([c0]), ring, u)
    want = _profile_mean(A, u)
    assert A_hex[0] == pytest.approx(want, rel=1e-11)
    x2, x4 = S.disk_moments(u)
    assert t4[0] == pytest.approx(c4 * rho ** 4 * x4, rel=1e-9)


def test_hexadecapole_means_equal_numpy_means():
    """The four-sample means are summed in the order numpy's mean sums four
    columns, so the features (and the gate calibrated on them) keep their
    values to the bit: checked against the expansion written with
    ndarray.mean over magnitudes from 1e-5 to 1e5."""
    rng = np.random.default_rng(7)
    x2, x4 = S.disk_moments(0.5)
    for trial in range(20):
        scale = 10.0 ** rng.integers(-5, 6)
        amp = scale * (1.0 + rng.random(512))
        ring = amp[:, None] * (1.0 + rng.normal(0.0, 1e-3, (512, 12)))
        half = ring[:, 0:4].mean(axis=1) - amp
        full = ring[:, 4:8].mean(axis=1) - amp
        diag = ring[:, 8:12].mean(axis=1) - amp
        a2 = (16.0 * half - full) / 3.0
        a4 = 0.5 * (full + diag) - a2
        A_hex, t4 = S.hexadecapole(amp, ring, 0.5)
        assert np.array_equal(A_hex.view(np.uint64), (amp + a2 * x2 + a4 * x4).view(np.uint64))
        assert np.array_equal(t4.view(np.uint64), (a4 * x4).view(np.uint64))


def _numpy_point_amp(lens, zeta):
    """Images from numpy's roots of the oracle's polynomial (numpy's
    companion-matrix eigenvalues, not Laguerre), kept when they satisfy the
    lens equation, and the sum of 1/|det J|."""
    from oracle import oracle as O
    a, b, th, m2, m3 = lens
    R = O.Oracle()
    zs = np.array([0.0, a, b * np.exp(1j * th)])
    ms = np.array([1 - m2 - m3, m2, m3])
    c = R.coeffs(a, b, th, m2, m3, zeta)
    roots = np.roots(c[::-1])
    amp = 0.0
    for z in roots:
        # numpy's roots of these polynomials satisfy the lens equation to
        # ~1e-11 where they are images and miss it by O(0.1) where not
        if abs(z - np.sum(ms / np.conj(z - zs)) - zeta) < 1e-8:
            s = np.sum(ms / (z - zs) ** 2)
            amp += 1.0 / abs(1.0 - abs(s) ** 2)
    return amp


def test_oracle_ring_amp_matches_independent_images(oracle):
    from triplelens_amd import workloads as W
    c = W.c1()
    lens = c.lens()
    idx = [0, 40, 120, 199]
    t = np.array(idx) / c.lc_len
    pos = (c.sx + (c.ex - c.sx) * t) + 1j * (c.sy + (c.ey - c.sy) * t)
    roots, _, _ = oracle.images(*lens, pos)
    ring, cross = oracle.ring_amp(*lens, c.rho, pos, roots)
    cc = math.sqrt(0.5)
    off = np.array([0.5, 0.5j, -0.5, -0.5j, 1, 1j, -1, -1j,
                    cc + cc * 1j, -cc + cc * 1j, -cc - cc * 1j, cc - cc * 1j])
    for i, p in enumerate(pos):
        want = np.array([_numpy_point_amp(lens, p + c.rho * o) for o in off])
        np.testing.assert_allclose(ring[i], want, rtol=1e-9)
        assert cross[i] == 0  # these positions are far from every caustic


def test_oracle_ring_crossings_at_a_caustic(oracle):
    """A source disk that straddles a caustic has samples on both sides:
    positions on a caustic sample of C1 count crossings, positions far from
    every caustic none."""
    from triplelens_amd import workloads as W
    c = W.c1()
    lens = c.lens()
    _, ca = oracle.ccc(*lens)
    on = ca.ravel()[::97][:20]
    roots, _, _ = oracle.images(*lens, on)
    _, cross = oracle.ring_amp(*lens, 1e-3, on, roots)
    assert np.mean(cross > 0) > 0.8
    far = np.array([3.0 + 3.0j, -3.0 - 2.0j])
    roots, _, _ = oracle.images(*lens, far)
    _, cross = oracle.ring_amp(*lens, 1e-3, far, roots)
    assert np.all(cross == 0)


def test_oracle_ring_ignores_a_flickering_faint_image(oracle):
    """Far from the caustics of the C2 geometry with seed 2 (hundreds of
    source radii), a faint image (magnification ~4e-6) next to a lens has a
    lens-equation residual at the 1e-6 acceptance threshold and is accepted
    at some disk samples but not at the centre: the image counts differ, the
    counts of significant images do not."""
    from triplelens_amd import workloads as W
    c = W.c2(seed=2, ca_fn=lambda *g: oracle.ccc(*g)[1])
    t = np.array([2, 4, 5, 7, 8]) / c.lc_len
    pos = (c.sx + (c.ex - c.sx) * t) + 1j * (c.sy + (c.ey - c.sy) * t)
    roots, _, nimg = oracle.images(*c.lens(), pos)
    ring, cross = oracle.ring_amp(*c.lens(), c.rho, pos, roots)
    cc = math.sqrt(0.5)
    off = np.array([0.5, 0.5j, -0.5, -0.5j, 1, 1j, -1, -1j,
                    cc + cc * 1j, -cc + cc * 1j, -cc - cc * 1j, cc - cc * 1j])
    flicker = 0
    for i, p in enumerate(pos):
        _, _, ns = oracle.images(*c.lens(), p + c.rho * off)
        flicker += np.count_nonzero(ns != nimg[i])
    assert flicker > 0
    assert np.all(cross == 0)


def test_constant_columns_keep_unit_scale(monkeypatch):
    """A feature that is constant up to rounding (log10 rho of a fixed-rho
    workload) must not be divided by its ~1e-16 deviation."""
    net = S.Net.__new__(S.Net)
    got = {}
    monkeypatch.setattr(S.Net, "set_scaling", lambda self, m, s: got.update(m=m, s=s))
    X = np.zeros((100, S.N_FEATURES))
    X[:, 0] = np.linspace(0, 1, 100)
    X[:, 5] = math.log10(5e-4) + np.where(np.arange(100) % 2, 4.4e-16, 0.0)
    net.fit_scaling(X, np.linspace(-1, 1, 100))
    assert got["s"][5] == 1.0 and got["s"][1] == 1.0
    assert got["s"][0] == pytest.approx(np.std(X[:, 0]))


def test_hexadecapole_feature_round_trip():
    """The compressed hexadecapole-term feature decodes to |a4 <r^4>| / A0."""
    c = math.sqrt(0.5)
    off = [(0.5, 0), (0, 0.5), (-0.5, 0), (0, -0.5), (1, 0), (0, 1), (-1, 0), (0, -1),
           (c, c), (-c, c), (-c, -c), (c, -c)]
    for c4 in (-3e-1, 1e-3, 2e1):
        A = lambda x, y: 2.0 + 0.3 * (x * x + y * y) + c4 * (x * x + y * y) ** 2
        ring = np.array([[A(x, y) for x, y in off]])
        _, t4 = S.hexadecapole(np.array([2.0]), ring, 0.5)
        X = S.features((1.0, 1.0, 1.0, 0.1, 0.1), 1e-3, 0.5, np.zeros((1, 2)), np.zeros((1, 2)),
                       np.array([1.0]), np.array([2.0]), np.array([1.0]), ring)
        assert S.hex_term(X)[0] == pytest.approx(abs(t4[0]) / 2.0, rel=1e-9)


def test_union_of_seed_sets():
    """Shot positions take the union of two seed searches: every pixel
    once, the first search's order first, at most 64."""
    from triplelens_amd.surrogate import union_seeds
    ns1 = np.array([2, 0, 64], np.int32)
    ns2 = np.array([3, 1, 1], np.int32)
    p1 = np.zeros((3, 64, 2), np.int32)
    p2 = np.zeros((3, 64, 2), np.int32)
    p1[0, :2] = [[5, 6], [7, 8]]
    p2[0, :3] = [[7, 8], [1, 2], [5, 6]]
    p2[1, 0] = [9, 9]
    p1[2] = np.stack([np.arange(64), np.arange(64)], 1)
    p2[2, 0] = [100, 100]
    ns, pix = union_seeds(ns1, p1, ns2, p2)
    assert ns.tolist() == [3, 1, 64]
    assert pix[0, :3].tolist() == [[5, 6], [7, 8], [1, 2]]
    assert pix[1, 0].tolist() == [9, 9]
    assert np.array_equal(pix[2], p1[2])


def test_extra_seed_searches_share_no_piece_boundary():
    """The extra seed searches of the shot positions use piece counts whose
    boundaries differ from the first search's everywhere but the start."""
    from triplelens_amd import surrogate as S
    sur = S.Surrogate(None, seed_piece=16)
    n = 4096
    p = sur.seed_parts(n)
    parts = [p] + sur.seed_parts_extra(n)
    assert len(set(parts)) == 4
    bounds = [set(n * q // k for q in range(1, k)) for k in parts]
    for i in range(4):
        for j in range(i + 1, 4):
            assert len(bounds[i] & bounds[j]) < 0.1 * len(bounds[i])
