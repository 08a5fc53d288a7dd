"""The drop-in library liblcirs.so on the CPU side: it loads, exports every
entry point include/tl_lcirs.h declares, builds the reference's grid and
critical curves on the host bit for bit, and reports errors through
tl_last_error instead of crashing.  No GPU work is started here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, curve_from_params, golden


@pytest.fixture(scope="module")
def lib():
    from triplelens_amd import build, engine
    if not os.path.exists(engine.LIB_PATH):
        build.build()
    return engine.load()


def test_exports_every_declared_symbol(lib):
    hdr = open(os.path.join(ROOT, "include", "tl_lcirs.h")).read()
    names = re.findall(r"^\s*(?:void\*|void|int|const char\*)\s+(\w+)\s*\(", hdr, re.M)
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n


def test_reference_abi_names_present(lib):
    # lcirs.cc:515-609 and the ctypes class LC_irs (ctypes_classes.py:128-193)
    for n in ["lcirs_new", "get_lc", "get_lc_irs", "copy_lc", "set_limb_darkening",
              "set_amoeba_printout"]:
        assert hasattr(lib, n)


def test_host_setup_matches_reference(lib, oracle):
    from triplelens_amd import engine as E
    for name in ["c1.npz", "c2_ppr20.npz"]:
        g = golden(name)
        c = curve_from_params(g["params"])
        lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        lc.set_limb_darkening("linear", c.u)
        assert np.array_equal(lc.grid()[:32], g["grid"])
    g = golden("c1.npz")
    c = curve_from_params(g["params"])
    cc, ca = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr).cc_ca()
    assert np.array_equal(cc, g["cc"]) and np.array_equal(ca, g["ca"])


def test_c3_host_setup_matches_oracle(lib, oracle):
    from triplelens_amd import engine as E, workloads as W
    for c in W.c3()[:24]:
        lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        lc.set_limb_darkening("linear", c.u)
        lo = oracle.new(*c.lens(), c.rho, c.lc_len, c.ppr, c.u)
        assert np.array_equal(lc.grid(), lo.grid())


def test_invalid_input_returns_null_with_error(lib):
    from triplelens_amd import engine as E
    with pytest.raises(E.EngineError, match="sourceSize"):
        E.LightCurve(1.0, 1.0, 1.0, 0.1, 0.1, 0.0, 10, 50)
    with pytest.raises(E.EngineError, match="lcLength"):
        E.LightCurve(1.0, 1.0, 1.0, 0.1, 0.0, 1e-3, 0, 50)


def test_binary_lens_grid_and_caustics(lib, oracle):
    """m3 = 0 builds the reference's binary lens (lcirs.cc:29-36): the
    critical curves of the light curve's triple-lens form with m3 = 0 and
    the grid from them, every value the oracle's to the bit."""
    from triplelens_amd import engine as E
    geo = (1.0, 0.9433981132056605, 1.0121970114513341, 0.2, 0.0)
    lc = E.LightCurve(*geo, 1e-3, 10, 50)
    cc, ca = lc.cc_ca()
    occ, oca = oracle.ccc(*geo)
    assert np.array_equal(cc, occ) and np.array_equal(ca, oca)
    o = oracle.new(*geo, 1e-3, 10, 50)
    assert np.array_equal(lc.grid()[:32], o.grid()[:32])


def test_limb_darkening_models(lib):
    from triplelens_amd import engine as E
    lc = E.LightCurve(1.0, 0.9, 1.0, 0.1, 0.1, 1e-3, 10, 50)
    g = lc.grid()
    assert (g[32], g[33]) == (0.6, 0.4)  # default profile (limbdarkeningmodel.cc:8-16)
    lc.set_limb_darkening("linear", 0.5)
    i0 = 1.0 / (1.0 - 0.5 / 3.0)
    g = lc.grid()
    assert (g[32], g[33]) == (i0 * 0.5, i0 * 0.5)
    assert g[7] == 1.0 / np.pi / 2500.0


def test_ctypes_class_mirrors_reference(lib):
    from triplelens_amd.ctypes_classes import LC_irs
    lc = LC_irs(1.0, 0.9433981132056605, 1.0121970114513341, 0.1, 0.1, 1e-3, 199, 50)
    lc.set_limb_darkening(b"linear", 0.0)
    lc.set_limb_darkening("default", 0.0)


def test_branch_free_complex_routines_match_general_ones():
    """The kernels' branch-free modulus, division and square root give the
    general routines' results to the bit (random normal, tiny, huge, zero,
    infinite and NaN operands; run on the host)."""
    import ctypes
    from triplelens_amd import engine as E
    lib = E.load()
    lib.tl_math_selftest.argtypes = [ctypes.c_int, ctypes.c_ulonglong]
    for seed in (1, 12345, 2 ** 40 + 7):
        assert lib.tl_math_selftest(300000, seed) == 0


def test_new_entry_points_reject_bad_arguments(lib):
    """The clone, batched seed search and fills of chosen positions check
    their arguments before any GPU work and report through tl_last_error."""
    import ctypes
    from triplelens_amd import engine as E
    L = E.load()
    L.lcirs_clone.restype = ctypes.c_void_p
    assert L.lcirs_clone(None) is None
    assert b"null" in L.tl_last_error()
    L.lcirs_batch_find_seeds.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    assert L.lcirs_batch_find_seeds(-1, None, None) == -1
    assert b"bad batch arguments" in L.tl_last_error()
    L.lcirs_fill_positions_seeds.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6
    assert L.lcirs_fill_positions_seeds(None, 1, None, None, None, None, None, None) == -1
    assert b"bad arguments" in L.tl_last_error()
    lc = E.LightCurve(1.0, 0.9, 1.0, 0.1, 0.1, 1e-3, 10, 50)
    with pytest.raises(E.EngineError, match="no seed search"):
        E.batch_fill_positions([lc], [np.array([0])])
    idx = np.array([0], np.int32)
    with pytest.raises(E.EngineError, match="no seed search"):
        lc.fill_positions(idx, (np.array([1], np.int32), np.zeros((1, 64, 2), np.int32)))
    with pytest.raises(E.EngineError, match="no seed search"):
        lc.seeds(idx)
    # a clone is a separate object with the original's set-up
    cl = lc.clone()
    assert cl.h != lc.h and np.array_equal(cl.grid(), lc.grid())
    # and a successful call clears the last error
    assert L.lcirs_clone(None) is None
    lc.grid()
    assert E.last_error() == ""


def test_host_image_polynomial_is_the_references(lib):
    """tl_image_poly on the host (the header the kernels compile) gives the
    reference's coefficients (imgpointcoeff.cc:223-447) to the bit."""
    from triplelens_amd import engine as E
    rows = golden("coeffs.npz")["rows"]
    for r in rows:
        a, b, th, m2, m3, zr, zi = r[:7]
        ref = r[7:].view(np.complex128)
        got = E.image_poly(a, b, th, m2, m3, [complex(zr, zi)])[0]
        assert np.array_equal(got.view(np.float64), ref.view(np.float64))


def test_host_image_polynomial_equals_oracle(lib, oracle):
    from triplelens_amd import engine as E
    rng = np.random.default_rng(19)
    for _ in range(500):
        a, b = np.exp(rng.uniform(np.log(0.3), np.log(3.0), 2))
        th = rng.uniform(0, 2 * np.pi)
        m2, m3 = np.exp(rng.uniform(np.log(1e-4), np.log(0.5), 2))
        z = rng.uniform(-2, 2, 8) + 1j * rng.uniform(-2, 2, 8)
        got = E.image_poly(a, b, th, m2, m3, z)
        for k in range(len(z)):
            want = oracle.coeffs(a, b, th, m2, m3, z[k])
            assert np.array_equal(got[k].view(np.float64), want.view(np.float64))


def test_slot_trace_arguments_are_checked(lib):
    from triplelens_amd import engine as E, workloads as W
    c = W.c1()
    lc = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
    lc.set_slot_trace(np.zeros(4, np.complex128))
    lc.set_slot_trace(None)
    lib.lcirs_set_slot_trace.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    assert lib.lcirs_set_slot_trace(lc.h, None, 5) == -1
    assert lib.lcirs_set_slot_trace(None, None, 0) == -1


def test_object_with_given_critical_curves(lib):
    """lcirs_new_ccc with a geometry's own critical curves builds the object
    lcirs_new builds (grid, caustic boxes, curves); null arrays are refused."""
    from triplelens_amd import engine as E, workloads as W
    for c in W.c3()[:6]:
        a = E.LightCurve(*c.lens(), c.rho, c.lc_len, c.ppr)
        cc, ca = a.cc_ca()
        b = E.LightCurve.with_critical_curves(*c.lens(), c.rho, c.lc_len, c.ppr, cc, ca)
        assert np.array_equal(a.grid(), b.grid())
        cc2, ca2 = b.cc_ca()
        assert np.array_equal(cc, cc2) and np.array_equal(ca, ca2)
    L = E.load()
    assert L.lcirs_new_ccc(1.0, 0.9, 1.0, 0.1, 0.1, 1e-3, 10, 50, None, None) is None
    assert b"null critical-curve" in L.tl_last_error()
