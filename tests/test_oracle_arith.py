"""The oracle's restatements of the complex-arithmetic library routines the
reference's path calls (glibc hypot and csqrt, libgcc complex division) agree
with the system's own routines bit for bit."""
import numpy as np


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    mant = rng.uniform(0.5, 1.0, size=(n, 4))
    expo = rng.integers(-40, 40, size=(n, 4))
    sign = rng.choice([-1.0, 1.0], size=(n, 4))
    v = sign * np.ldexp(mant, expo)
    v[::7, 1] = 0.0        # pure real
    v[::11, 0] = 0.0       # pure imaginary
    return v


def test_hypot_matches_libm(oracle):
    v = _inputs(20000, 1)
    lib = oracle.lib
    bad = sum(lib.tlo_hypot(x, y) != lib.tlo_libm_hypot(x, y) for x, y, _, _ in v)
    assert bad == 0


def test_csqrt_matches_libm(oracle):
    v = _inputs(20000, 2)
    a = np.zeros(2)
    b = np.zeros(2)
    lib = oracle.lib
    bad = 0
    for x, y, _, _ in v:
        lib.tlo_csqrt(x, y, a.ctypes.data)
        lib.tlo_libm_csqrt(x, y, b.ctypes.data)
        bad += not (np.array_equal(a, b) and np.signbit(a[1]) == np.signbit(b[1]))
    assert bad == 0


def test_complex_division_matches_libgcc(oracle):
    v = _inputs(20000, 3)
    a = np.zeros(2)
    b = np.zeros(2)
    lib = oracle.lib
    bad = 0
    for x, y, c, d in v:
        lib.tlo_cdiv(x, y, c, d, a.ctypes.data)
        lib.tlo_libm_cdiv(x, y, c, d, b.ctypes.data)
        bad += not np.array_equal(a, b)
    assert bad == 0


def test_division_by_real_is_smith_with_zero_imaginary(oracle):
    # double / complex promotes the scalar to (t, 0): the form the image test
    # m_k / conj(z - z_k) takes (imgpoint.cc:74)
    a = np.zeros(2)
    b = np.zeros(2)
    for t, c, d in [(0.8, 0.3, -0.7), (1e-3, 2.5, 1e-9), (0.1, -1e-12, 4.0)]:
        oracle.lib.tlo_cdiv(t, 0.0, c, d, a.ctypes.data)
        oracle.lib.tlo_libm_cdiv(t, 0.0, c, d, b.ctypes.data)
        assert np.array_equal(a, b)
