"""Degree-10 image polynomial of the oracle.

The oracle evaluates the reference's generated expressions in their
operation order (getCoeffsOptNoZ + getCoeffsOptJustZ, imgpointcoeff.cc:
223-447, summed as imgpoint.cc:172-176 does), so its coefficients are the
reference's to the bit.  Pinned by (1) the reference's coefficients (fixture
coeffs.npz, and live comparisons with the reference build where it is
present) bit for bit, (2) an independent derivation of the same polynomial
(Q-product form, double-double, rounded once), which exact rational
arithmetic pins in turn: the two agree within the reference's rounding
error, which a dropped or misplaced term would exceed, and (3) the invariant
P(z) = 0 at lens-equation solutions (images built by mapping random z
forward)."""
import math
from fractions import Fraction as F

import numpy as np

from conftest import golden


class _C:
    __slots__ = ("r", "i")

    def __init__(self, r, i=0):
        self.r, self.i = F(r), F(i)

    def __add__(self, o):
        return _C(self.r + o.r, self.i + o.i)

    def __sub__(self, o):
        return _C(self.r - o.r, self.i - o.i)

    def __mul__(self, o):
        return _C(self.r * o.r - self.i * o.i, self.r * o.i + self.i * o.r)


def _pm(p, q):
    out = [_C(0) for _ in range(len(p) + len(q) - 1)]
    for i, x in enumerate(p):
        for j, y in enumerate(q):
            out[i + j] = out[i + j] + x * y
    return out


def _exact(a, b, th, m2, m3, zeta):
    """Coefficients in exact rational arithmetic from the same double inputs
    (lens frame of lens.cc:35-40, m1 = 1 - m2 - m3 in double)."""
    m = [F(1.0 - m2 - m3), F(m2), F(m3)]
    s, c = math.sin(th), math.cos(th)
    z = [_C(0.0), _C(a), _C(b * c, b * s)]
    f = [[_C(-zk.r, -zk.i), _C(1)] for zk in z]
    d12, d13, d23 = _pm(f[0], f[1]), _pm(f[0], f[2]), _pm(f[1], f[2])
    D = _pm(d12, f[2])
    S = [_C(m[0]) * d23[i] + _C(m[1]) * d13[i] + _C(m[2]) * d12[i] for i in range(3)]
    zc = _C(zeta.real, -zeta.imag)
    Q = [[(zc - _C(zk.r, -zk.i)) * D[i] + (S[i] if i < 3 else _C(0)) for i in range(4)] for zk in z]
    q12, q13, q23 = _pm(Q[0], Q[1]), _pm(Q[0], Q[2]), _pm(Q[1], Q[2])
    left = _pm([_C(-zeta.real, -zeta.imag), _C(1)], _pm(q12, Q[2]))
    mix = [_C(m[0]) * q23[i] + _C(m[1]) * q13[i] + _C(m[2]) * q12[i] for i in range(7)]
    right = _pm(D, mix)
    return np.array([complex(float(x.r), float(x.i)) for x in
                     [(right[i] if i < 10 else _C(0)) - left[i] for i in range(11)]])


def test_exact_derivation_is_correctly_rounded(oracle):
    cases = [(1.0, 0.9433981132056605, 1.0121970114513341, 0.1, 0.1, 0.3 + 0.1j),
             (0.56, 0.69, 5.03, 0.0277, 0.0017, 0.05 - 0.02j),
             (2.2, 0.41, 2.5, 3e-4, 0.2, -0.7 + 0.4j)]
    for a, b, th, m2, m3, z in cases:
        got = oracle.coeffs_exact(a, b, th, m2, m3, z)
        assert np.array_equal(got, _exact(a, b, th, m2, m3, z))


def _random_cases(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        a, b = np.exp(rng.uniform(np.log(0.3), np.log(3.0), 2))
        th = rng.uniform(0, 2 * np.pi)
        m2, m3 = np.exp(rng.uniform(np.log(1e-4), np.log(0.5), 2))
        if rng.random() < 0.1:
            m3 = 0.0  # the binary lens keeps the same polynomial with m3 = 0
        z = complex(rng.uniform(-2, 2), rng.uniform(-2, 2))
        yield float(a), float(b), float(th), float(m2), float(m3), z


def test_reference_rounding_agrees_with_the_exact_polynomial(oracle):
    """Same polynomial: the reference-order evaluation lies within its
    rounding error (1e-13 of the largest coefficient) of the exact values."""
    for a, b, th, m2, m3, z in _random_cases(2000, 7):
        got = oracle.coeffs(a, b, th, m2, m3, z)
        ex = oracle.coeffs_exact(a, b, th, m2, m3, z)
        assert np.max(np.abs(got - ex)) <= 1e-13 * np.max(np.abs(ex))


def test_polynomial_vanishes_at_images(oracle):
    rng = np.random.default_rng(7)
    for _ in range(40):
        a, b = rng.uniform(0.3, 3.0, 2)
        th = rng.uniform(0, 2 * np.pi)
        m2, m3 = np.exp(rng.uniform(np.log(1e-4), np.log(0.3), 2))
        m1 = 1.0 - m2 - m3
        zk = np.array([0.0, a, b * np.exp(1j * th)])
        z = complex(rng.uniform(-2, 2), rng.uniform(-2, 2))
        zeta = z - np.sum(np.array([m1, m2, m3]) / np.conj(z - zk))
        for c in (oracle.coeffs(a, b, th, m2, m3, zeta), oracle.coeffs_exact(a, b, th, m2, m3, zeta)):
            val = np.polyval(c[::-1], z)
            scale = np.sum(np.abs(c) * np.abs(z) ** np.arange(11))
            assert abs(val) <= 1e-13 * scale


def test_equal_to_reference_coefficients(oracle):
    rows = golden("coeffs.npz")["rows"]
    for r in rows:
        a, b, th, m2, m3, zr, zi = r[:7]
        ref = r[7:].view(np.complex128)
        got = oracle.coeffs(a, b, th, m2, m3, complex(zr, zi))
        assert np.array_equal(got.view(np.float64), ref.view(np.float64))


def test_equal_to_live_reference(oracle, reference):
    for a, b, th, m2, m3, z in _random_cases(3000, 11):
        got = oracle.coeffs(a, b, th, m2, m3, z)
        ref = reference.coeffs(a, b, th, m2, m3, z)
        assert np.array_equal(got.view(np.float64), ref.view(np.float64)), (a, b, th, m2, m3, z)
