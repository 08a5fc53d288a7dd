"""tl_mlp_predict_noalloc, the surrogate inference that allocates nothing:
its results equal tl_mlp_predict's for the same rows, on two calls in a row
(the second with fewer rows), and more rows than its scratch array holds are
rejected.  The test allocates the device buffers it passes in (through the
HIP runtime the library uses); the entry point itself allocates nothing."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


@pytest.fixture(scope="module")
def hip():
    from triplelens_amd import engine
    engine.load()  # liblcirs.so and, with it, its HIP runtime
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(_P), ctypes.c_size_t]
    lib.hipFree.argtypes = [_P]
    lib.hipMemcpy.argtypes = [_P, _P, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _dev(hip, nbytes):
    p = _P()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def test_noalloc_inference_equals_predict(hip):
    from triplelens_amd import surrogate as S
    lib = S._lib()
    lib.tl_mlp_predict_noalloc.argtypes = [_P, ctypes.c_int, _P, _P]
    net = S.Net(seed=5)
    rng = np.random.default_rng(3)
    net.set_scaling(rng.normal(size=S.N_FEATURES).astype(np.float32),
                    rng.uniform(0.5, 2.0, S.N_FEATURES).astype(np.float32))
    H2D, D2H = 1, 2
    cap = 4096
    dx = _dev(hip, cap * S.N_FEATURES * 4)
    dout = _dev(hip, cap * 2 * 4)
    try:
        for n in (3000, 517):  # two calls in a row, the second smaller
            x = rng.normal(size=(n, S.N_FEATURES)).astype(np.float32)
            want = np.zeros((n, 2), np.float32)
            assert lib.tl_mlp_predict(net.h, n, x.ctypes.data, want.ctypes.data) == 0
            assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
            assert lib.tl_mlp_predict_noalloc(net.h, n, dx, dout) == 0, \
                lib.tl_mlp_last_error().decode()
            got = np.zeros((n, 2), np.float32)
            assert hip.hipMemcpy(got.ctypes.data, dout, got.nbytes, D2H) == 0
            assert np.array_equal(got, want), n
        assert lib.tl_mlp_predict_noalloc(net.h, cap + 1, dx, dout) == -1
        assert b"at most 4096" in lib.tl_mlp_last_error()
        assert lib.tl_mlp_predict_noalloc(net.h, 0, None, None) == 0
    finally:
        hip.hipFree(dx)
        hip.hipFree(dout)
