"""Python classes over liblcirs.so with the names, arguments and behaviour of
the reference's ctypes wrapper (Examples/ctypes_classes.py:128-193, and
LC_cuda at :196-272).

``LC_irs`` is the reference class: same constructor arguments (the eighth,
named ``img_plane_size`` there, is pointsPerRadius, lcirs.cc:524), same
methods, output written into a caller-allocated float64 array of length
``lc_len`` by ``copy_lc``.  The only differences are that the library is
found next to this package instead of at '../bin/liblcirs.so' and that
errors raise instead of being printed.
"""
from __future__ import annotations

import ctypes

import numpy as np
from numpy.ctypeslib import ndpointer

from . import engine as _engine


class LC_irs(object):
    def __init__(self, a, b, th, m2, m3, source_size, lc_len, img_plane_size):
        _engine.load()  # raises when the library has not been built
        # a handle of its own, declared as the reference declares it
        self.lib_lc = lib = ctypes.CDLL(_engine.LIB_PATH)
        D = ctypes.c_double
        lib.lcirs_new.argtypes = [D, D, D, D, D, D, ctypes.c_uint, ctypes.c_long]
        lib.lcirs_new.restype = ctypes.c_void_p
        lib.get_lc.argtypes = [ctypes.c_void_p, D, D, D, D]
        lib.get_lc_irs.argtypes = [ctypes.c_void_p, D, D, D, D]
        lib.copy_lc.argtypes = [ctypes.c_void_p, ndpointer(dtype=np.double, flags="C_CONTIGUOUS")]
        lib.set_limb_darkening.argtypes = [ctypes.c_void_p, ctypes.c_char_p, D]
        lib.set_amoeba_printout.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
        lib.lcirs_delete.argtypes = [ctypes.c_void_p]
        lib.lcirs_copy_ray_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.lcirs_copy_lc_full.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.obj = lib.lcirs_new(a, b, th, m2, m3, source_size, lc_len, img_plane_size)
        if not self.obj:
            raise _engine.EngineError(_engine.last_error() or "lcirs_new failed")
        self.lc_len = int(lc_len)

    def __del__(self):
        obj = getattr(self, "obj", None)
        if obj:
            self.lib_lc.lcirs_delete(obj)
            self.obj = None

    def get_lc(self, iniX, iniY, finX, finY):
        self.lib_lc.get_lc(self.obj, iniX, iniY, finX, finY)
        _engine._check()

    def get_lc_irs(self, iniX, iniY, finX, finY):
        self.lib_lc.get_lc_irs(self.obj, iniX, iniY, finX, finY)
        _engine._check()

    def copy_lc(self, lc_vec):
        self.lib_lc.copy_lc(self.obj, lc_vec)

    def set_limb_darkening(self, model_type, model_parameter):
        if isinstance(model_type, str):
            model_type = model_type.encode()
        self.lib_lc.set_limb_darkening(self.obj, model_type, model_parameter)
        _engine._check()

    def set_amoeba_printout(self, amoeba_filename, par_filename):
        if isinstance(amoeba_filename, str):
            amoeba_filename = amoeba_filename.encode()
        if isinstance(par_filename, str):
            par_filename = par_filename.encode()
        self.lib_lc.set_amoeba_printout(self.obj, amoeba_filename, par_filename)
        _engine._check()

    # additions
    def ray_counts(self):
        out = np.zeros(self.lc_len + 1, np.uint64)
        _engine._check(self.lib_lc.lcirs_copy_ray_counts(self.obj, out.ctypes.data))
        return out

    def lc_full(self):
        out = np.zeros(self.lc_len + 1)
        _engine._check(self.lib_lc.lcirs_copy_lc_full(self.obj, out.ctypes.data))
        return out


class LC_cuda(object):
    """The reference's LC_cuda (Examples/ctypes_classes.py:196-272) on this
    engine: lccuda_new / get_lc_cuda / set_limb_darkening_cuda, and get_lc /
    copy_lc shared with LC_irs.  get_lc_cuda gives the exact FP64 curve."""

    def __init__(self, a, b, th, m2, m3, source_size, lc_len, img_plane_size):
        _engine.load()
        self.lib_cuda = lib = ctypes.CDLL(_engine.LIB_PATH)
        D = ctypes.c_double
        lib.lccuda_new.argtypes = [D, D, D, D, D, D, ctypes.c_uint, ctypes.c_long]
        lib.lccuda_new.restype = ctypes.c_void_p
        lib.get_lc.argtypes = [ctypes.c_void_p, D, D, D, D]
        lib.get_lc_cuda.argtypes = [ctypes.c_void_p, D, D, D, D]
        lib.copy_lc.argtypes = [ctypes.c_void_p, ndpointer(dtype=np.double, flags="C_CONTIGUOUS")]
        lib.set_limb_darkening_cuda.argtypes = [ctypes.c_void_p, ctypes.c_char_p, D]
        lib.lcirs_delete.argtypes = [ctypes.c_void_p]
        self.obj = lib.lccuda_new(a, b, th, m2, m3, source_size, lc_len, img_plane_size)
        if not self.obj:
            raise _engine.EngineError(_engine.last_error() or "lccuda_new failed")
        self.lc_len = int(lc_len)

    def __del__(self):
        obj = getattr(self, "obj", None)
        if obj:
            self.lib_cuda.lcirs_delete(obj)
            self.obj = None

    def get_lc(self, iniX, iniY, finX, finY):
        self.lib_cuda.get_lc(self.obj, iniX, iniY, finX, finY)
        _engine._check()

    def get_lc_cuda(self, iniX, iniY, finX, finY):
        self.lib_cuda.get_lc_cuda(self.obj, iniX, iniY, finX, finY)
        _engine._check()

    def copy_lc(self, lc_vec):
        self.lib_cuda.copy_lc(self.obj, lc_vec)

    def set_limb_darkening(self, model_type, model_parameter):
        if isinstance(model_type, str):
            model_type = model_type.encode()
        self.lib_cuda.set_limb_darkening_cuda(self.obj, model_type, model_parameter)
        _engine._check()
