"""ctypes loaders for the two CPU checkers.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module; the product
(``triplelens_amd``) never imports it.

* :class:`Oracle` wraps ``_build/libtloracle.so``, the C restatement of the
  reference's LightCurveIRS path (``tl_oracle.c``).
* :class:`Reference` wraps ``_ref/libtlref.so``, the reference's own sources
  compiled by ``Makefile`` (present only where ``/root/reference`` was
  available at build time).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libtloracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtlref.so")
REF_FAST_SO = os.path.join(HERE, "_ref", "libtlref_fast.so")

_D = ctypes.c_double
_P = ctypes.c_void_p


def build(ref: bool = False) -> None:
    """Compile the C restatement (and, when asked, the reference library)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Oracle:
    """The C restatement.  One instance per process is enough."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.tlo_lcirs_new.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long]
        lib.tlo_lcirs_new.restype = _P
        lib.tlo_lcirs_delete.argtypes = [_P]
        lib.tlo_set_limb_darkening.argtypes = [_P, ctypes.c_int, _D]
        lib.tlo_set_slot_trace.argtypes = [_P, _P, ctypes.c_long]
        lib.tlo_frac_oob.argtypes = [_P, _D, _D, _D, _D, _P]
        lib.tlo_grid.argtypes = [_P, _P]
        lib.tlo_get_lc_irs.argtypes = [_P, _D, _D, _D, _D, _P, _P, _P]
        lib.tlo_seeds.argtypes = [_P, _D, _D, _D, _D, ctypes.c_int, _P, _P, _P]
        lib.tlo_fill_point.argtypes = [_P, _D, _D, ctypes.c_int, _P, _P, _P, _P]
        lib.tlo_ccc.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P]
        lib.tlo_coeffs.argtypes = [_D] * 7 + [_P]
        lib.tlo_coeffs_exact.argtypes = [_D] * 7 + [_P]
        lib.tlo_laguerre_solve.argtypes = [ctypes.c_int, _P, _P, _P]
        lib.tlo_laguerre_one.argtypes = [ctypes.c_int, _P, _P]
        lib.tlo_images.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P, _P, _P]
        lib.tlo_irs.argtypes = [_P, _D, _D, _D, _D]
        lib.tlo_irs.restype = _D
        lib.tlo_point_amp.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P]
        lib.tlo_ring_amp.argtypes = [_D] * 6 + [ctypes.c_int, _P, _P, _P, _P]
        lib.tlo_hypot.argtypes = [_D, _D]
        lib.tlo_hypot.restype = _D
        lib.tlo_csqrt.argtypes = [_D, _D, _P]
        lib.tlo_cdiv.argtypes = [_D, _D, _D, _D, _P]
        lib.tlo_batch.argtypes = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_int, _P, _P]
        lib.tlo_libm_hypot.argtypes = [_D, _D]
        lib.tlo_libm_hypot.restype = _D
        lib.tlo_libm_csqrt.argtypes = [_D, _D, _P]
        lib.tlo_libm_cdiv.argtypes = [_D, _D, _D, _D, _P]
        lib.tlo_seeds_margin.argtypes = [_P, _D, _D, _D, _D, ctypes.c_int, _P, _P, _P, _P]
        lib.tlo_fill_stats.argtypes = [_P]
        lib.tlo_chain_states.argtypes = [_P, _D, _D, _D, _D, ctypes.c_uint, ctypes.c_uint,
                                         ctypes.c_int, _P]
        lib.tlo_last_segments.argtypes = [_P, _P, ctypes.c_long]
        lib.tlo_last_segments.restype = ctypes.c_long
        lib.tlo_hit_mask.argtypes = [_P, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                     ctypes.c_long, _D, _D, _P]
        self.lib = lib

    # -- light curve object -------------------------------------------------
    def new(self, a, b, th, m2, m3, rho, lc_len, ppr, ld_v=None):
        h = self.lib.tlo_lcirs_new(a, b, th, m2, m3, rho, lc_len, ppr)
        if not h:
            raise ValueError("tlo_lcirs_new rejected its arguments")
        if ld_v is not None:
            self.lib.tlo_set_limb_darkening(h, 1, float(ld_v))
        return LC(self, h, lc_len)

    def ccc(self, a, b, th, m2, m3, length=500):
        cc = np.zeros(6 * length, np.complex128)
        ca = np.zeros(6 * length, np.complex128)
        self.lib.tlo_ccc(a, b, th, m2, m3, length, _ptr(cc), _ptr(ca))
        return cc.reshape(6, length), ca.reshape(6, length)

    def coeffs(self, a, b, th, m2, m3, zeta):
        out = np.zeros(11, np.complex128)
        self.lib.tlo_coeffs(a, b, th, m2, m3, zeta.real, zeta.imag, _ptr(out))
        return out

    def coeffs_exact(self, a, b, th, m2, m3, zeta):
        """The same polynomial from an independent derivation, each
        coefficient the exact value rounded once."""
        out = np.zeros(11, np.complex128)
        self.lib.tlo_coeffs_exact(a, b, th, m2, m3, zeta.real, zeta.imag, _ptr(out))
        return out

    def laguerre_solve(self, coeffs, init=None):
        c = np.ascontiguousarray(coeffs, np.complex128)
        deg = len(c) - 1
        r = np.zeros(deg, np.complex128)
        ini = None if init is None else np.ascontiguousarray(init, np.complex128)
        if ini is not None and len(ini) != deg + 1:
            raise ValueError("init needs deg+1 values")
        self.lib.tlo_laguerre_solve(deg, _ptr(c), None if ini is None else _ptr(ini), _ptr(r))
        return r

    def laguerre_one(self, coeffs, x0):
        c = np.ascontiguousarray(coeffs, np.complex128)
        x = np.array([x0], np.complex128)
        ok = self.lib.tlo_laguerre_one(len(c) - 1, _ptr(c), _ptr(x))
        return bool(ok), complex(x[0])

    def images(self, a, b, th, m2, m3, pos):
        p = np.ascontiguousarray(pos, np.complex128)
        n = len(p)
        roots = np.zeros((n, 10), np.complex128)
        is_img = np.zeros((n, 10), np.int32)
        n_img = np.zeros(n, np.int32)
        self.lib.tlo_images(a, b, th, m2, m3, n, _ptr(p), _ptr(roots), _ptr(is_img), _ptr(n_img))
        return roots, is_img.astype(bool), n_img

    def point_amp(self, a, b, th, m2, m3, pos):
        p = np.ascontiguousarray(pos, np.complex128)
        out = np.zeros(len(p))
        self.lib.tlo_point_amp(a, b, th, m2, m3, len(p), _ptr(p), _ptr(out))
        return out

    def ring_amp(self, a, b, th, m2, m3, rho, pos, roots):
        """Point magnification at the 12 disk samples around each position
        (complex pos (n,)), solved from the given roots (n, 10): (n, 12), and
        the samples per position whose count of significant images (each at
        least 1e-4 of the magnification) differs from the centre's (n,)."""
        p = np.ascontiguousarray(pos, np.complex128)
        r = np.ascontiguousarray(roots, np.complex128)
        if r.shape != (len(p), 10):
            raise ValueError("roots must be (n, 10)")
        out = np.zeros((len(p), 12))
        cross = np.zeros(len(p), np.int32)
        self.lib.tlo_ring_amp(a, b, th, m2, m3, rho, len(p), _ptr(p), _ptr(r), _ptr(out),
                              _ptr(cross))
        return out, cross

    def batch(self, params, traj, lens, nthreads):
        """Independent curves over host threads.  params: (n, 8) rows of
        a, b, th, m2, m3, rho, ppr, v; traj: (n, 4); lens: lcLength per curve."""
        params = np.ascontiguousarray(params, np.float64)
        traj = np.ascontiguousarray(traj, np.float64)
        lens = np.ascontiguousarray(lens, np.uint32)
        off = np.zeros(len(lens), np.uint64)
        off[1:] = np.cumsum(lens.astype(np.uint64) + 1)[:-1]
        total = int((lens.astype(np.uint64) + 1).sum())
        amp = np.zeros(total)
        rays = np.zeros(total, np.uint64)
        rc = self.lib.tlo_batch(len(lens), _ptr(params), _ptr(traj), _ptr(lens), _ptr(off),
                                nthreads, _ptr(amp), _ptr(rays))
        if rc:
            raise RuntimeError(f"tlo_batch failed ({rc})")
        return amp, rays, off


class LC:
    """One restated LightCurveIRS object."""

    def __init__(self, owner: Oracle, handle, lc_len):
        self._o = owner
        self.h = handle
        self.lc_len = lc_len

    def __del__(self):
        if getattr(self, "h", None):
            self._o.lib.tlo_lcirs_delete(self.h)
            self.h = None

    def grid(self):
        g = np.zeros(36)
        self._o.lib.tlo_grid(self.h, _ptr(g))
        return g

    def lc_irs(self, sx, sy, ex, ey):
        n = self.lc_len + 1
        amp = np.zeros(n)
        rays = np.zeros(n, np.uint64)
        edge = np.zeros(n, np.uint8)
        rc = self._o.lib.tlo_get_lc_irs(self.h, sx, sy, ex, ey, _ptr(amp), _ptr(rays), _ptr(edge))
        if rc:
            raise RuntimeError(f"tlo_get_lc_irs failed ({rc})")
        return amp, rays, edge.astype(bool)

    def frac_oob(self, sx, sy, ex, ey):
        """Per point: Laguerre kick steps that read the reference's 8-entry
        fractional-step table past its end (the result there is undefined in
        the reference)."""
        out = np.zeros(self.lc_len + 1, np.int32)
        rc = self._o.lib.tlo_frac_oob(self.h, sx, sy, ex, ey, _ptr(out))
        if rc:
            raise RuntimeError(f"tlo_frac_oob failed ({rc})")
        return out

    def set_slot_trace(self, slots):
        """Start the chained solve of getImages call k (from the next light
        curve on) with tempRoots[10] = slots[k] (complex) instead of 0; None
        switches it off."""
        if slots is None:
            self._slots = None
            self._o.lib.tlo_set_slot_trace(self.h, None, 0)
            return
        self._slots = np.ascontiguousarray(slots, np.complex128)
        self._o.lib.tlo_set_slot_trace(self.h, _ptr(self._slots), len(self._slots))

    def seeds(self, sx, sy, ex, ey, max_seeds=64):
        n = self.lc_len + 1
        xy = np.zeros((n, max_seeds, 2))
        ns = np.zeros(n, np.int32)
        ni = np.zeros(n, np.int32)
        self._o.lib.tlo_seeds(self.h, sx, sy, ex, ey, max_seeds, _ptr(xy), _ptr(ns), _ptr(ni))
        return xy, ns, ni

    def seeds_margin(self, sx, sy, ex, ey, max_seeds=64):
        """Seeds as seeds(), plus per point the distance in decades between
        the closest root residual and the 1e-6 image threshold."""
        n = self.lc_len + 1
        xy = np.zeros((n, max_seeds, 2))
        ns = np.zeros(n, np.int32)
        ni = np.zeros(n, np.int32)
        mg = np.zeros(n)
        self._o.lib.tlo_seeds_margin(self.h, sx, sy, ex, ey, max_seeds, _ptr(xy), _ptr(ns),
                                     _ptr(ni), _ptr(mg))
        return xy, ns, ni, mg

    def fill_point(self, px, py, seeds_xy):
        s = np.ascontiguousarray(seeds_xy, np.float64).reshape(-1, 2)
        rays = np.zeros(1, np.uint64)
        ssum = np.zeros(1)
        edge = np.zeros(1, np.int32)
        rc = self._o.lib.tlo_fill_point(self.h, px, py, len(s), _ptr(s), _ptr(rays), _ptr(ssum),
                                        _ptr(edge))
        if rc:
            raise RuntimeError(f"tlo_fill_point failed ({rc})")
        return int(rays[0]), float(ssum[0]), bool(edge[0])

    def chain_states(self, sx, sy, ex, ey, i0, i1, fresh):
        """Root-chain state after each point i0..i1: (n, 11) complex roots and
        (n,) have-roots flags."""
        n = i1 - i0 + 1
        out = np.zeros((n, 23))
        if self._o.lib.tlo_chain_states(self.h, sx, sy, ex, ey, i0, i1, int(fresh), _ptr(out)):
            raise ValueError("bad chain range")
        return out[:, :22].copy().view(np.complex128), out[:, 22].astype(int)

    def last_segments(self):
        """The amoeba of the last fill_point: (n, 3) rows of ny, xleft, xright,
        sorted by row and column."""
        n = self._o.lib.tlo_last_segments(self.h, None, 0)
        out = np.zeros((max(n, 0), 3), np.int64)
        if n > 0:
            self._o.lib.tlo_last_segments(self.h, _ptr(out), n)
        return out[np.lexsort((out[:, 1], out[:, 0]))] if n > 0 else out

    def irs(self, x, y, sx, sy):
        return self._o.lib.tlo_irs(self.h, x, y, sx, sy)

    def hit_mask(self, x0, x1, y0, y1, sx, sy):
        """Boolean hit mask of pixels [y0..y1] x [x0..x1] for source (sx, sy)."""
        out = np.zeros((y1 - y0 + 1, x1 - x0 + 1), np.uint8)
        if self._o.lib.tlo_hit_mask(self.h, x0, x1, y0, y1, sx, sy, _ptr(out)):
            raise ValueError("bad hit-mask box")
        return out.astype(bool)


class Reference:
    """The reference's own LightCurveIRS path (compiled from its sources)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = ctypes.CDLL(path)
        lib.tlref_grid.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long, ctypes.c_int, _D, _P]
        lib.tlref_lc_irs.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long, ctypes.c_int,
                                                _D, _D, _D, _D, _D, _P, _P]
        lib.tlref_seeds.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long, _D, _D, _D, _D,
                                               ctypes.c_int, _P, _P, _P]
        lib.tlref_ccc.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P]
        lib.tlref_images.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P, _P, _P]
        lib.tlref_point_lc.argtypes = [_D] * 5 + [ctypes.c_uint, _D, _D, _D, _D, _P]
        lib.tlref_point_amp.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P]
        lib.tlref_laguerre_solve.argtypes = [ctypes.c_int, _P, _P, _P]
        lib.tlref_coeffs.argtypes = [_D] * 7 + [_P]
        lib.tlref_lc_irs_slots.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long, ctypes.c_int,
                                                      _D, _D, _D, _D, _D, _P, _P, _P, ctypes.c_long]
        lib.tlref_lc_irs_slots.restype = ctypes.c_long
        lib.tlref_lc_irs_print.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long, ctypes.c_int,
                                                      _D, _D, _D, _D, _D, ctypes.c_char_p,
                                                      ctypes.c_char_p]
        self.lib = lib

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def grid(self, a, b, th, m2, m3, rho, lc_len, ppr, ld_v=None):
        g = np.zeros(32)
        self.lib.tlref_grid(a, b, th, m2, m3, rho, lc_len, ppr,
                            0 if ld_v is None else 1, 0.0 if ld_v is None else ld_v, _ptr(g))
        return g

    def lc_irs_print(self, a, b, th, m2, m3, rho, lc_len, ppr, ld_v, sx, sy, ex, ey,
                     amoeba_prefix, par_file):
        """getLCIRS with the reference's amoeba and parameter print-out on."""
        self.lib.tlref_lc_irs_print(a, b, th, m2, m3, rho, lc_len, ppr, 1, ld_v, sx, sy, ex, ey,
                                    amoeba_prefix.encode(), par_file.encode())

    def lc_irs(self, a, b, th, m2, m3, rho, lc_len, ppr, ld_v, sx, sy, ex, ey):
        n = lc_len + 1
        amp = np.zeros(n)
        rays = np.zeros(n, np.uint64)
        parsed = self.lib.tlref_lc_irs(a, b, th, m2, m3, rho, lc_len, ppr,
                                       0 if ld_v is None else 1, 0.0 if ld_v is None else ld_v,
                                       sx, sy, ex, ey, _ptr(amp), _ptr(rays))
        if parsed != n:
            raise RuntimeError(f"parsed {parsed} ray counts, expected {n}")
        return amp, rays

    def lc_irs_slots(self, a, b, th, m2, m3, rho, lc_len, ppr, ld_v, sx, sy, ex, ey,
                     max_slots=1 << 22):
        """lc_irs, plus the value the chained solve of every getImages call
        read one element past the previous roots (complex; NaN where there
        was no such element)."""
        n = lc_len + 1
        amp = np.zeros(n)
        rays = np.zeros(n, np.uint64)
        slots = np.zeros(2 * max_slots)
        calls = self.lib.tlref_lc_irs_slots(a, b, th, m2, m3, rho, lc_len, ppr,
                                            0 if ld_v is None else 1,
                                            0.0 if ld_v is None else ld_v, sx, sy, ex, ey,
                                            _ptr(amp), _ptr(rays), _ptr(slots), max_slots)
        if calls < 0 or calls > max_slots:
            raise RuntimeError(f"slot recording failed ({calls} calls)")
        return amp, rays, slots[:2 * calls].view(np.complex128).copy()

    def seeds(self, a, b, th, m2, m3, rho, lc_len, ppr, sx, sy, ex, ey, max_seeds=64):
        n = lc_len + 1
        xy = np.zeros((n, max_seeds, 2))
        ns = np.zeros(n, np.int32)
        ni = np.zeros(n, np.int32)
        self.lib.tlref_seeds(a, b, th, m2, m3, rho, lc_len, ppr, sx, sy, ex, ey, max_seeds,
                             _ptr(xy), _ptr(ns), _ptr(ni))
        return xy, ns, ni

    def ccc(self, a, b, th, m2, m3, length=500):
        cc = np.zeros(6 * length, np.complex128)
        ca = np.zeros(6 * length, np.complex128)
        self.lib.tlref_ccc(a, b, th, m2, m3, length, _ptr(cc), _ptr(ca))
        return cc.reshape(6, length), ca.reshape(6, length)

    def coeffs(self, a, b, th, m2, m3, zeta):
        out = np.zeros(11, np.complex128)
        self.lib.tlref_coeffs(a, b, th, m2, m3, zeta.real, zeta.imag, _ptr(out))
        return out

    def images(self, a, b, th, m2, m3, pos):
        p = np.ascontiguousarray(pos, np.complex128)
        n = len(p)
        roots = np.zeros((n, 10), np.complex128)
        is_img = np.zeros((n, 10), np.int32)
        n_img = np.zeros(n, np.int32)
        self.lib.tlref_images(a, b, th, m2, m3, n, _ptr(p), _ptr(roots), _ptr(is_img), _ptr(n_img))
        return roots, is_img.astype(bool), n_img

    def point_amp(self, a, b, th, m2, m3, pos):
        p = np.ascontiguousarray(pos, np.complex128)
        out = np.zeros(len(p))
        self.lib.tlref_point_amp(a, b, th, m2, m3, len(p), _ptr(p), _ptr(out))
        return out

    def laguerre_solve(self, coeffs, init=None):
        c = np.ascontiguousarray(coeffs, np.complex128)
        deg = len(c) - 1
        r = np.zeros(deg, np.complex128)
        ini = None if init is None else np.ascontiguousarray(init, np.complex128)
        self.lib.tlref_laguerre_solve(deg, _ptr(c), None if ini is None else _ptr(ini), _ptr(r))
        return r
