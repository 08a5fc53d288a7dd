"""ctypes binding of liblcirs.so (the C ABI in include/tl_lcirs.h).

The library is the product: a HIP engine for gfx950.  Loading fails loudly if
it has not been built, and every GPU call raises when the engine reports an
error; there is no CPU path behind it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

LIB_PATH = _build.LIB
MAX_SEEDS = 64

FLAG_EDGE = 1
FLAG_SEED_OVERFLOW = 2
FLAG_HASH_FULL = 4
FLAG_LABEL_FULL = 8
FLAG_QUEUE_FULL = 16
FLAG_TIMEOUT = 32
FLAG_ROW_OVERFLOW = 64
FLAG_KICK_TABLE = 128   # the reference's result at the point is undefined (see tl_lcirs.h)
# flags that report on a point without making its result incomplete
FLAGS_INFO = FLAG_EDGE | FLAG_KICK_TABLE

_D = ctypes.c_double
_P = ctypes.c_void_p
_lib = None


class EngineError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads liblcirs.so (once) and declares its functions."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"{path} is missing: build it with triplelens_amd/build.py")
    lib = ctypes.CDLL(path)
    lib.lcirs_new.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long]
    lib.lcirs_new.restype = _P
    lib.lcirs_new_ccc.argtypes = [_D] * 6 + [ctypes.c_uint, ctypes.c_long, _P, _P]
    lib.lcirs_new_ccc.restype = _P
    lib.lcirs_delete.argtypes = [_P]
    lib.lcirs_clone.argtypes = [_P]
    lib.lcirs_clone.restype = _P
    lib.lcirs_batch_find_seeds.argtypes = [ctypes.c_int, _P, _P]
    for f in ("get_lc", "get_lc_irs"):
        getattr(lib, f).argtypes = [_P, _D, _D, _D, _D]
        getattr(lib, f).restype = None
    lib.copy_lc.argtypes = [_P, _P]
    lib.set_limb_darkening.argtypes = [_P, ctypes.c_char_p, _D]
    lib.set_amoeba_printout.argtypes = [_P, ctypes.c_char_p, ctypes.c_char_p]
    for f in ("lcirs_copy_lc_full", "lcirs_copy_ray_counts", "lcirs_copy_flags",
              "lcirs_get_grid"):
        getattr(lib, f).argtypes = [_P, _P]
        getattr(lib, f).restype = ctypes.c_int
    lib.lcirs_copy_cc_ca.argtypes = [_P, _P, _P]
    lib.lcirs_copy_seeds.argtypes = [_P, _P, _P]
    lib.lcirs_batch_irs.argtypes = [ctypes.c_int, _P, _P]
    lib.lcirs_fill_seeds.argtypes = [_P, ctypes.c_int, _P, _P, _P, _P, _P, _P]
    lib.tl_last_stats.argtypes = [_P]
    lib.tl_set_option.argtypes = [ctypes.c_char_p, _D]
    lib.tl_last_error.restype = ctypes.c_char_p
    lib.tl_device_info.argtypes = [ctypes.c_char_p, ctypes.c_int, _P]
    _lib = lib
    return lib


def last_error() -> str:
    return load().tl_last_error().decode()


def _check(rc: int = 0):
    err = last_error()
    if rc != 0 or err:
        raise EngineError(err or f"engine call failed ({rc})")


def set_option(name: str, value: float) -> None:
    if load().tl_set_option(name.encode(), float(value)) != 0:
        raise ValueError(f"unknown engine option {name}")


def last_stats() -> dict:
    out = np.zeros(21)
    load().tl_last_stats(out.ctypes.data)
    keys = ["tasks", "evals", "runs", "dups", "seed_ms", "fill_ms", "fill_launches", "total_ms",
            "busy_ms", "insert_ms", "idle_polls", "waves", "scouts", "scout_runs", "trial_ms",
            "ring_ms", "solves", "spec_pieces", "spec_repairs", "spec_fallback", "graph_builds"]
    return dict(zip(keys, out.tolist()))


def device_info():
    name = ctypes.create_string_buffer(128)
    cus = ctypes.c_int(0)
    _check(load().tl_device_info(name, 128, ctypes.byref(cus)))
    return name.value.decode(), cus.value


class LightCurve:
    """One LightCurveIRS object of the engine."""

    def __init__(self, a, b, th, m2, m3, rho, lc_len, ppr, _handle=None):
        lib = load()
        self._lib = lib
        self.h = _handle if _handle else lib.lcirs_new(a, b, th, m2, m3, rho, lc_len, ppr)
        if not self.h:
            raise EngineError(last_error() or "lcirs_new failed")
        self.lc_len = int(lc_len)

    @classmethod
    def with_critical_curves(cls, a, b, th, m2, m3, rho, lc_len, ppr, cc, ca):
        """lcirs_new_ccc: the object with the critical curves and caustics
        given (one geometry's (6, 500) complex arrays, as ccc_batch gives
        them) instead of computed on the host."""
        lib = load()
        cc = np.ascontiguousarray(cc, np.complex128).reshape(6, 500)
        ca = np.ascontiguousarray(ca, np.complex128).reshape(6, 500)
        h = lib.lcirs_new_ccc(a, b, th, m2, m3, rho, lc_len, ppr, cc.ctypes.data, ca.ctypes.data)
        if not h:
            raise EngineError(last_error() or "lcirs_new_ccc failed")
        return cls(0, 0, 0, 0, 0, 0, lc_len, 0, _handle=h)

    def clone(self):
        """A new object with this one's set-up, options and root chain
        (lcirs_clone), without its results."""
        h = self._lib.lcirs_clone(self.h)
        if not h:
            raise EngineError(last_error() or "lcirs_clone failed")
        return LightCurve(0, 0, 0, 0, 0, 0, self.lc_len, 0, _handle=h)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self._lib.lcirs_delete(h)
            self.h = None

    def set_limb_darkening(self, model: str, v: float):
        self._lib.set_limb_darkening(self.h, model.encode(), float(v))
        _check()

    def set_slot_trace(self, slots):
        """lcirs_set_slot_trace: getImages call k of this object's chain (from
        the next light curve on) starts its chained solve with tempRoots[10]
        = slots[k] (complex) instead of 0; None or empty clears it."""
        self._lib.lcirs_set_slot_trace.argtypes = [_P, _P, ctypes.c_long]
        if slots is None or len(slots) == 0:
            _check(self._lib.lcirs_set_slot_trace(self.h, None, 0))
            return
        xy = np.ascontiguousarray(slots, np.complex128)
        _check(self._lib.lcirs_set_slot_trace(self.h, xy.ctypes.data, len(xy)))

    def lc_irs(self, sx, sy, ex, ey):
        """Runs get_lc_irs; returns (magnification, rays, flags), lcLength+1 each."""
        self._lib.get_lc_irs(self.h, sx, sy, ex, ey)
        _check()
        return self.results()

    def results(self):
        n = self.lc_len + 1
        amp = np.zeros(n)
        rays = np.zeros(n, np.uint64)
        flags = np.zeros(n, np.uint32)
        _check(self._lib.lcirs_copy_lc_full(self.h, amp.ctypes.data))
        _check(self._lib.lcirs_copy_ray_counts(self.h, rays.ctypes.data))
        _check(self._lib.lcirs_copy_flags(self.h, flags.ctypes.data))
        return amp, rays, flags

    def flags(self):
        """Status flags of every position of the last seed search or light
        curve (TL_FLAG_* bits)."""
        flags = np.zeros(self.lc_len + 1, np.uint32)
        _check(self._lib.lcirs_copy_flags(self.h, flags.ctypes.data))
        return flags

    def point_amp(self):
        """Point-source magnification, weighted image centroid and signed
        caustic distance at the positions of the last seed search:
        (n,), (n, 2), (n,) with n = lcLength+1."""
        n = self.lc_len + 1
        amp = np.zeros(n)
        cen = np.zeros((n, 2))
        caus = np.zeros(n)
        self._lib.lcirs_copy_point_amp.argtypes = [_P, _P, _P, _P, _P]
        _check(self._lib.lcirs_copy_point_amp(self.h, amp.ctypes.data, cen.ctypes.data,
                                              caus.ctypes.data, None))
        return amp, cen, caus

    def pair_share(self):
        """Share of the point magnification in the two brightest images, at
        the positions of the last seed search."""
        out = np.zeros(self.lc_len + 1)
        self._lib.lcirs_copy_point_amp.argtypes = [_P, _P, _P, _P, _P]
        _check(self._lib.lcirs_copy_point_amp(self.h, None, None, None, out.ctypes.data))
        return out

    def set_ring(self, on: bool = True):
        """With on, later seed searches also compute the point magnification
        at 12 disk samples around each position (lcirs_set_ring)."""
        self._lib.lcirs_set_ring.argtypes = [_P, ctypes.c_int]
        _check(self._lib.lcirs_set_ring(self.h, 1 if on else 0))

    def set_seed_split(self, parts: int):
        """Run the seed search in `parts` independent pieces
        (lcirs_set_seed_split); 1 restores the single chain."""
        self._lib.lcirs_set_seed_split.argtypes = [_P, ctypes.c_int]
        _check(self._lib.lcirs_set_seed_split(self.h, int(parts)))

    def ring_amp(self, with_roots: bool = True):
        """Disk-sample magnifications (n, 12), each position's own roots
        (n, 10; None without with_roots) and the number of samples across a
        caustic from the centre (n,) of the last seed search, n =
        lcLength+1."""
        n = self.lc_len + 1
        ring = np.zeros((n, 12))
        roots = np.zeros((n, 10), np.complex128) if with_roots else None
        cross = np.zeros(n, np.int32)
        self._lib.lcirs_copy_ring_amp.argtypes = [_P, _P, _P, _P]
        _check(self._lib.lcirs_copy_ring_amp(self.h, ring.ctypes.data,
                                             roots.ctypes.data if with_roots else None,
                                             cross.ctypes.data))
        return ring, roots, cross

    def find_seeds(self, sx, sy, ex, ey):
        """The seed search alone (no fill); see point_amp and seeds."""
        self._lib.lcirs_find_seeds.argtypes = [_P, _D, _D, _D, _D]
        _check(self._lib.lcirs_find_seeds(self.h, sx, sy, ex, ey))

    def point_lc(self, sx, sy, ex, ey):
        self._lib.get_lc(self.h, sx, sy, ex, ey)
        _check()
        out = np.zeros(self.lc_len)
        self._lib.copy_lc(self.h, out.ctypes.data)
        return out

    def copy_lc(self):
        out = np.zeros(self.lc_len)
        self._lib.copy_lc(self.h, out.ctypes.data)
        return out

    def grid(self):
        g = np.zeros(36)
        _check(self._lib.lcirs_get_grid(self.h, g.ctypes.data))
        return g

    def cc_ca(self):
        cc = np.zeros(6 * 500, np.complex128)
        ca = np.zeros(6 * 500, np.complex128)
        _check(self._lib.lcirs_copy_cc_ca(self.h, cc.ctypes.data, ca.ctypes.data))
        return cc.reshape(6, 500), ca.reshape(6, 500)

    def seeds(self, idx=None):
        """Seed counts (n,) and pixels (n, 64, 2) of the last seed search, at
        every position, or at positions idx only (lcirs_copy_seeds_at)."""
        if idx is not None:
            idx = np.ascontiguousarray(idx, np.int32)
            ns = np.empty(len(idx), np.int32)
            pix = np.empty((len(idx), MAX_SEEDS, 2), np.int32)
            self._lib.lcirs_copy_seeds_at.argtypes = [_P, ctypes.c_int, _P, _P, _P]
            _check(self._lib.lcirs_copy_seeds_at(self.h, len(idx), idx.ctypes.data,
                                                 ns.ctypes.data, pix.ctypes.data))
            return ns, pix
        n = self.lc_len + 1
        ns = np.zeros(n, np.int32)
        pix = np.zeros((n, MAX_SEEDS, 2), np.int32)
        _check(self._lib.lcirs_copy_seeds(self.h, ns.ctypes.data, pix.ctypes.data))
        return ns, pix

    def seeds_union(self, others, idx):
        """Seed counts and pixels at positions idx of the union of this
        object's seeds and those of `others` (lcirs_copy_seeds_union): this
        object's seeds first, then each other's new ones, at most 64."""
        idx = np.ascontiguousarray(idx, np.int32)
        ns = np.empty(len(idx), np.int32)
        pix = np.empty((len(idx), MAX_SEEDS, 2), np.int32)
        arr = (ctypes.c_void_p * max(1, len(others)))(*[o.h for o in others])
        self._lib.lcirs_copy_seeds_union.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, _P, _P, _P]
        _check(self._lib.lcirs_copy_seeds_union(self.h, len(others), arr, len(idx),
                                                idx.ctypes.data, ns.ctypes.data, pix.ctypes.data))
        return ns, pix

    def fill_seeds(self, pos, nseeds, pix):
        """GPU fill of given pixel seeds at given source positions."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 2)
        n = len(pos)
        nseeds = np.ascontiguousarray(nseeds, np.int32)
        pix = np.ascontiguousarray(pix, np.int32).reshape(n, MAX_SEEDS, 2)
        rays = np.zeros(n, np.uint64)
        amp = np.zeros(n)
        flags = np.zeros(n, np.uint32)
        _check(self._lib.lcirs_fill_seeds(self.h, n, pos.ctypes.data, nseeds.ctypes.data,
                                          pix.ctypes.data, rays.ctypes.data, amp.ctypes.data,
                                          flags.ctypes.data))
        return amp, rays, flags

    def fill_positions(self, idx, seeds=None):
        """GPU fill of positions idx (indices 0..lcLength) of the last seed
        search (lcirs_fill_positions): amp, rays, flags.  With seeds =
        (nseeds, pix) for those positions ((n,), (n, 64, 2)), the given seeds
        replace the search's (lcirs_fill_positions_seeds)."""
        idx = np.ascontiguousarray(idx, np.int32)
        n = len(idx)
        rays = np.zeros(n, np.uint64)
        amp = np.zeros(n)
        flags = np.zeros(n, np.uint32)
        if n == 0:
            return amp, rays, flags
        if seeds is None:
            self._lib.lcirs_fill_positions.argtypes = [_P, ctypes.c_int, _P, _P, _P, _P]
            _check(self._lib.lcirs_fill_positions(self.h, n, idx.ctypes.data, rays.ctypes.data,
                                                  amp.ctypes.data, flags.ctypes.data))
            return amp, rays, flags
        ns = np.ascontiguousarray(seeds[0], np.int32)
        pix = np.ascontiguousarray(seeds[1], np.int32).reshape(n, MAX_SEEDS, 2)
        if len(ns) != n:
            raise ValueError("one seed count per position")
        self._lib.lcirs_fill_positions_seeds.argtypes = [_P, ctypes.c_int] + [_P] * 6
        _check(self._lib.lcirs_fill_positions_seeds(self.h, n, idx.ctypes.data, ns.ctypes.data,
                                                    pix.ctypes.data, rays.ctypes.data,
                                                    amp.ctypes.data, flags.ctypes.data))
        return amp, rays, flags


def image_poly(a, b, th, m2, m3, zeta, on_device=False):
    """tl_image_poly: the degree-10 image polynomial (coefficients of z^0..z^10)
    at each source position of ``zeta``, as the reference rounds it; on the
    host or with the engine's GPU kernel.  Returns an (n, 11) complex array."""
    lib = load()
    z = np.ascontiguousarray(np.atleast_1d(zeta), np.complex128)
    out = np.zeros((len(z), 11), np.complex128)
    lib.tl_image_poly.argtypes = [_D] * 5 + [ctypes.c_int, _P, _P, ctypes.c_int]
    _check(lib.tl_image_poly(a, b, th, m2, m3, len(z), z.ctypes.data, out.ctypes.data,
                             1 if on_device else 0))
    return out


def ccc_batch(params):
    """Critical curves and caustics of many geometries on the GPU.
    params: (n, 5) rows of a, b, th, m2, m3.  Returns cc, ca as (n, 6, 500)."""
    lib = load()
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 5)
    n = len(p)
    cc = np.zeros((n, 6, 500), np.complex128)
    ca = np.zeros((n, 6, 500), np.complex128)
    lib.tl_ccc_batch.argtypes = [ctypes.c_int, _P, _P, _P]
    _check(lib.tl_ccc_batch(n, p.ctypes.data, cc.ctypes.data, ca.ctypes.data))
    return cc, ca


def batch_find_seeds(curves, trajs):
    """The seed searches of many LightCurve objects in one engine run
    (lcirs_batch_find_seeds); trajs: (n, 4)."""
    lib = load()
    n = len(curves)
    arr = (ctypes.c_void_p * n)(*[c.h for c in curves])
    t = np.ascontiguousarray(trajs, np.float64).reshape(n, 4)
    _check(lib.lcirs_batch_find_seeds(n, arr, t.ctypes.data))


def batch_fill_positions(curves, idx_list, seeds_list=None):
    """lcirs_batch_fill_positions: positions idx_list[k] of curves[k]'s last
    seed search, all in one engine run, with (unless seeds_list is None) the
    seeds seeds_list[k] = (nseeds, pix); returns [(amp, rays, flags)]."""
    lib = load()
    n = len(curves)
    counts = np.array([len(i) for i in idx_list], np.int32)
    tot = int(counts.sum())
    idx = np.ascontiguousarray(np.concatenate([np.asarray(i, np.int32) for i in idx_list])
                               if n else np.zeros(0, np.int32), np.int32)
    rays = np.zeros(tot, np.uint64)
    amp = np.zeros(tot)
    flags = np.zeros(tot, np.uint32)
    ns_p = pix_p = None
    if seeds_list is not None:
        ns = np.ascontiguousarray(np.concatenate([s[0] for s in seeds_list]), np.int32)
        pix = np.ascontiguousarray(np.concatenate([np.asarray(s[1]).reshape(-1, MAX_SEEDS, 2)
                                                   for s in seeds_list]), np.int32)
        if len(ns) != tot:
            raise ValueError("one seed count per position")
        ns_p, pix_p = ns.ctypes.data, pix.ctypes.data
    arr = (ctypes.c_void_p * max(1, n))(*[c.h for c in curves])
    lib.lcirs_batch_fill_positions.argtypes = [ctypes.c_int] + [_P] * 8
    _check(lib.lcirs_batch_fill_positions(n, arr, counts.ctypes.data, idx.ctypes.data, ns_p,
                                          pix_p, rays.ctypes.data, amp.ctypes.data,
                                          flags.ctypes.data))
    out, o = [], 0
    for k in range(n):
        m = int(counts[k])
        out.append((amp[o:o + m], rays[o:o + m], flags[o:o + m]))
        o += m
    return out


def batch_irs(curves, trajs):
    """One engine run over many LightCurve objects; trajs: (n, 4)."""
    lib = load()
    n = len(curves)
    arr = (ctypes.c_void_p * n)(*[c.h for c in curves])
    t = np.ascontiguousarray(trajs, np.float64).reshape(n, 4)
    _check(lib.lcirs_batch_irs(n, arr, t.ctypes.data))
    return [c.results() for c in curves]
