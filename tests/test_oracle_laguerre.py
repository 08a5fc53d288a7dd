"""Laguerre root finder of the oracle (src/Laguerre/laguerre.cc).

Pinned by the reference's own unit tests (tests/laguerretest.cc:32-142, same
polynomials and tolerance eps = 1e-10) and, bit for bit, by roots the
reference's solveRoots produced for random polynomials (fixture
laguerre.npz, fresh and from initial roots)."""
import numpy as np

from conftest import golden

EPS = 1.0e-10


def test_second_order_known_answer(oracle):
    # (x+1)^2 (laguerretest.cc:32-48)
    r = oracle.laguerre_solve(np.array([1, 2, 1], complex))
    assert len(r) == 2
    for z in r:
        assert abs(z.real - (-1.0)) <= 4 * np.spacing(1.0)  # EXPECT_DOUBLE_EQ
        assert abs(z) < 1.0 + EPS


def test_third_order_known_answer(oracle):
    # (x-1)(x-2)(x-3) (laguerretest.cc:50-83)
    r = np.sort_complex(oracle.laguerre_solve(np.array([-6, 11, -6, 1], complex)))
    for z, e in zip(sorted(r, key=lambda z: z.real), [1, 2, 3]):
        assert abs(z - e) < EPS and abs(z.imag) < EPS


def test_fourth_order_known_answer(oracle):
    # (x-1)(x-2)(x-3)(x-4) (laguerretest.cc:85-118)
    r = oracle.laguerre_solve(np.array([24, -50, 35, -10, 1], complex))
    for z, e in zip(sorted(r, key=lambda z: z.real), [1, 2, 3, 4]):
        assert abs(z - e) < EPS and abs(z.imag) < EPS


def test_polish_from_perturbed_roots(oracle):
    # laguerretest.cc:121-142, with fixed instead of random phases
    c = np.array([24, -50, 35, -10, 1], complex)
    for root in [1, 2, 3, 4]:
        for ph in np.linspace(0, 2 * np.pi, 16, endpoint=False):
            ok, z = oracle.laguerre_one(c, root + 1e-2 * np.exp(1j * ph))
            assert ok and abs(z - root) < EPS


def test_matches_reference_solve_bit_for_bit(oracle):
    g = golden("laguerre.npz")
    oc = ii = orr = 0
    for deg, has_init in zip(g["deg"], g["has_init"]):
        c = g["coeffs"][oc:oc + deg + 1]
        init = g["init"][ii:ii + deg + 1] if has_init else None
        want = g["roots"][orr:orr + deg]
        got = oracle.laguerre_solve(c, init)
        assert np.array_equal(got.view(np.float64), want.view(np.float64))
        oc += deg + 1
        ii += deg + 1
        orr += deg
