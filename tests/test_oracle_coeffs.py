"""Degree-10 image polynomial of the oracle.

The oracle derives the polynomial itself and evaluates it in double-double,
so each coefficient is the exact value for the double inputs, rounded once.
Pinned by (1) exact rational arithmetic on a few cases, (2) the invariant
P(z) = 0 at lens-equation solutions (images built by mapping random z
forward), (3) the reference's coefficients (imgpointcoeff.cc:223-447, fixture
coeffs.npz), which agree to the reference's own rounding error."""
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


def test_coefficients_are_correctly_rounded(oracle):
    cases = [(1.0, 0.9433981132056605, 1.0121970114513341, 0.1, 0.1, 0.3 + 0.1j),
             (0.56, 0.69, 5.03, 0.0277, 0.0017, 0.05 - 0.02j),
             (2.2, 0.41, 2.5, 3e-4, 0.2, -0.7 + 0.4j)]
    for a, b, th, m2, m3, z in cases:
        got = oracle.coeffs(a, b, th, m2, m3, z)
        assert np.array_equal(got, _exact(a, b, th, m2, m3, z))


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
        c = oracle.coeffs(a, b, th, m2, m3, zeta)
        val = np.polyval(c[::-1], z)
        scale = np.sum(np.abs(c) * np.abs(z) ** np.arange(11))
        assert abs(val) <= 1e-13 * scale


def test_agree_with_reference_coefficients(oracle):
    rows = golden("coeffs.npz")["rows"]
    for r in rows:
        a, b, th, m2, m3, zr, zi = r[:7]
        ref = r[7:].view(np.complex128)
        got = oracle.coeffs(a, b, th, m2, m3, complex(zr, zi))
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
