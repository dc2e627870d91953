"""ctypes loader for oracle/_ref/libgsref.so (the reference's own profile numerics).

Test infrastructure only (see ``oracle/__init__.py``).  ``available()`` is False when the
library has not been built (it is built in this repository's development container from the
reference sources by ``make -C oracle``; the built file travels with the tree).
"""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref', 'libgsref.so')
_lib = None


def available():
    return os.path.exists(_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_PATH)
        d = ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.gsref_sersic_info.argtypes = [d, d, dp, dp]
        L.gsref_sersic_kvalue.argtypes = [d, d, dp, dp, dp, ctypes.c_int]
        for name, nargs in (('gsref_sersic_hlr', 2), ('gsref_sersic_integrated_flux', 2),
                            ('gsref_sersic_truncated_scale', 3),
                            ('gsref_moffat_scale_radius_from_hlr', 3),
                            ('gsref_bessel_k', 2), ('gsref_bessel_j', 2), ('gsref_bessel_y', 2),
                            ('gsref_gamma_p', 2), ('gsref_hankel_sersic', 4),
                            ('gsref_hankel_kolmogorov', 3), ('gsref_expd', 1)):
            f = getattr(L, name)
            f.argtypes = [d] * nargs
            f.restype = d
        L.gsref_bessel_root.argtypes = [d, ctypes.c_int]
        L.gsref_bessel_root.restype = d
        L.gsref_kolmogorov_info.argtypes = [dp, dp]
        L.gsref_hankel_moffat_trunc.argtypes = [d] * 5
        L.gsref_hankel_moffat_trunc.restype = d
        ip = ctypes.POINTER(ctypes.c_int)
        L.gsref_kvalue_range_1d.argtypes = [ctypes.c_int, d, d, d, d, d, ip, dp]
        L.gsref_kvalue_range_1d.restype = None
        L.gsref_kvalue_range_2d.argtypes = [ctypes.c_int, d, d, d, d, d, d, ip]
        L.gsref_kvalue_range_2d.restype = None
        _lib = L
    return _lib


def _gsp_array(gsp):
    if gsp is None:
        return None
    a = np.array([gsp.minimum_fft_size, gsp.maximum_fft_size, gsp.folding_threshold,
                  gsp.stepk_minimum_hlr, gsp.maxk_threshold, gsp.kvalue_accuracy,
                  gsp.xvalue_accuracy, gsp.table_spacing, gsp.realspace_relerr,
                  gsp.realspace_abserr, gsp.integration_relerr, gsp.integration_abserr,
                  gsp.shoot_accuracy], dtype=float)
    return a


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def sersic_info(n, trunc=0., gsp=None):
    """(maxK, stepK, HLR, XNorm) of the reference SersicInfo, in units of r0."""
    g = _gsp_array(gsp)
    out = np.zeros(4)
    if lib().gsref_sersic_info(n, trunc, _ptr(g), _ptr(out)) != 0:
        raise RuntimeError("reference SersicInfo failed")
    return tuple(out)


def sersic_kvalue(n, ksq, trunc=0., gsp=None):
    g = _gsp_array(gsp)
    ksq = np.ascontiguousarray(ksq, dtype=float)
    out = np.zeros_like(ksq)
    if lib().gsref_sersic_kvalue(n, trunc, _ptr(g), _ptr(ksq), _ptr(out), len(ksq)) != 0:
        raise RuntimeError("reference SersicInfo failed")
    return out


def kolmogorov_info(gsp=None):
    g = _gsp_array(gsp)
    out = np.zeros(2)
    if lib().gsref_kolmogorov_info(_ptr(g), _ptr(out)) != 0:
        raise RuntimeError("reference KolmogorovInfo failed")
    return tuple(out)


def __getattr__(name):
    # sersic_hlr, bessel_k, gamma_p, ... -> gsref_<name>
    if name.startswith('__'):
        raise AttributeError(name)
    return getattr(lib(), 'gsref_' + name)


def kvalue_range_1d(m, kmax, ksqmax, kx0, dkx, ky):
    """The reference's GetKValueRange1d (src/SBProfile.cpp:508-530): (i1, i2)."""
    out = np.zeros(2, dtype=np.int32)
    kysq = np.zeros(1)
    lib().gsref_kvalue_range_1d(int(m), kmax, ksqmax, kx0, dkx, ky,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(kysq))
    return int(out[0]), int(out[1])


def kvalue_range_2d(m, kmax, ksqmax, kx0, dkx, ky0, dky):
    """The reference's GetKValueRange2d (src/SBProfile.cpp:532-588): (i1, i2)."""
    out = np.zeros(2, dtype=np.int32)
    lib().gsref_kvalue_range_2d(int(m), kmax, ksqmax, kx0, dkx, ky0, dky,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return int(out[0]), int(out[1])
