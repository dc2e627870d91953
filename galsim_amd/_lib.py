"""ctypes binding of libgsmi.so (the HIP kernels and the host profile numerics).

The library is built in-tree by ``galsim_amd.csrc.build``.  There is no fallback: if it is
missing, every drawing call raises ``GalSimAMDError`` naming the library.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgsmi.so")

OP_DTYPE = np.dtype([('code', np.int32), ('arg', np.int32), ('a', np.float64),
                     ('b', np.float64), ('c', np.float64)], align=True)
LEAF_DTYPE = np.dtype([('type', np.int32), ('separable', np.int32), ('kind', np.int32),
                       ('table', np.int32), ('ntab', np.int32), ('cliprows', np.int32),
                       ('jac', np.float64, (4,)), ('scale', np.float64), ('scale2', np.float64),
                       ('flux', np.float64), ('p', np.float64, (6,))], align=True)
STAMP_DTYPE = np.dtype([('N', np.int32), ('Nk', np.int32), ('nx', np.int32), ('ny', np.int32),
                        ('x0', np.int32), ('y0', np.int32), ('op_begin', np.int32),
                        ('op_count', np.int32), ('leaf_begin', np.int32),
                        ('axis_count', np.int32), ('out_offset', np.int64),
                        ('scratch_offset', np.int64), ('axis_offset', np.int64),
                        ('dk', np.float64), ('status', np.int32), ('col_stride', np.int32)],
                       align=True)
assert OP_DTYPE.itemsize == 32 and LEAF_DTYPE.itemsize == 128 and STAMP_DTYPE.itemsize == 80
STAMP_NONFINITE = 1      # device: a non-finite output pixel (gsmi_stamp.status)
STACK_DEPTH = {'f64': 6, 'f32': 4}
GRAD_MAXLEAF = 4
GRAD_PER_LEAF = 7
GRAD_FIELDS = 31
ADJ_COMPLEX = 1
ADJ_SHAPE = 2

OP_LEAF, OP_MUL, OP_ADD, OP_SCALE, OP_PHASE, OP_INV = 1, 2, 3, 4, 5, 6
LEAF_EXPONENTIAL, LEAF_GAUSSIAN, LEAF_MOFFAT, LEAF_KOLMOGOROV, LEAF_SERSIC, LEAF_BOX = range(1, 7)
LEAF_KCLIP = 7
LEAF_SPERGEL = 8
LEAF_AIRY = 9
LEAF_INCLINED_EXP = 10
LEAF_INCLINED_VERT = 11
LEAF_TOPHAT = 12
LEAF_DELTA = 13
LEAF_INTERP_IMAGE = 14

# GSMI_LEAF_INTERP_IMAGE parameter block (include/gsmi.h GSMI_INTERP_*)
INTERP_N2, INTERP_KSCALE, INTERP_MAXK1, INTERP_MAXK1S, INTERP_KXR = range(5)
INTERP_I1, INTERP_I2, INTERP_J1, INTERP_J2 = 5, 6, 7, 8
INTERP_XCODE, INTERP_XN, INTERP_XUMAX, INTERP_XUTAB, INTERP_XUTABN = 9, 10, 11, 12, 13
INTERP_KCODE, INTERP_KN, INTERP_KDC, INTERP_KK = 14, 15, 16, 17
INTERP_KIMAGE, INTERP_XMIN, INTERP_YMIN, INTERP_BLOCK = 22, 23, 24, 25
INTERP_MAX_TAPS = 17

ERRORS = {-1: "invalid argument", -2: "parameter out of range", -3: "solve/table failure",
          -4: "HIP runtime failure", -5: "unsupported FFT size"}


WORK_DTYPE = np.dtype([('stamp', np.int32), ('first', np.int32), ('count', np.int32),
                       ('per_group', np.int32)])


class Schedule(ctypes.Structure):
    """gsmi_schedule: the two work lists of a batch and their launch shapes."""
    _fields_ = [('fold', ctypes.c_void_p), ('rows', ctypes.c_void_p),
                ('nfold', ctypes.c_int32), ('nrows', ctypes.c_int32),
                ('fold_grid', ctypes.c_int32), ('rows_grid', ctypes.c_int32),
                ('fold_lds', ctypes.c_int32), ('rows_lds', ctypes.c_int32),
                ('program', ctypes.c_int32)]


# Program shapes with specialised fold kernels (gsmi_schedule.program): post-order ops
# ((code, arg), ...) and the leaf type of each slot.
STATIC_PROGRAMS = {
    1: (((1, 0), (1, 1), (3, 0), (1, 2), (2, 0)), (1, 5, 4)),   # (Exp + Sersic) x Kolmogorov
    2: (((1, 0), (1, 1), (2, 0)), (1, 3)),                     # Exponential x Moffat
    3: (((1, 0), (1, 1), (2, 0)), (5, 4)),                     # Sersic x Kolmogorov
    4: (((1, 0), (1, 1), (2, 0)), (1, 4)),                     # Exponential x Kolmogorov
}


class GalSimAMDError(RuntimeError):
    """Raised for failures inside the native library (mirrors GalSimError)."""


class GalSimAMDRangeError(GalSimAMDError, ValueError):
    """A parameter outside the range the reference accepts (mirrors GalSimRangeError)."""


class GalSimAMDValueError(GalSimAMDError, ValueError):
    """An argument of the wrong shape or value (mirrors GalSimValueError)."""


class GalSimAMDBoundsError(GalSimAMDError, ValueError):
    """Image bounds that do not include a required point (mirrors GalSimBoundsError)."""


class GalSimAMDUndefinedBoundsError(GalSimAMDError):
    """An operation on an image with undefined bounds (mirrors GalSimUndefinedBoundsError)."""


_lib = None

# name -> (restype, argtypes); every symbol declared in include/gsmi.h
_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_SIGS = {
    'gsmi_last_error': (ctypes.c_char_p, []),
    'gsmi_version': (_I, []),
    'gsmi_good_fft_size': (_I, [_I]),
    'gsmi_sersic_info': (_I, [_D, _D, _P, _P, _P]),
    'gsmi_sersic_table': (_I, [_D, _D, _P, _P, _P, _P, _I]),
    'gsmi_sersic_hlr': (_I, [_D, _D, _P]),
    'gsmi_sersic_integrated_flux': (_I, [_D, _D, _P]),
    'gsmi_sersic_truncated_scale': (_I, [_D, _D, _D, _P]),
    'gsmi_kolmogorov_info': (_I, [_P, _P]),
    'gsmi_moffat_info': (_I, [_D, _D, _D, _P, _P]),
    'gsmi_moffat_scale_radius_from_hlr': (_I, [_D, _D, _P]),
    'gsmi_moffat_scale_radius_from_hlr_trunc': (_I, [_D, _D, _D, _P]),
    'gsmi_moffat_trunc_info': (_I, [_D, _D, _D, _P, _P]),
    'gsmi_moffat_trunc_table': (_I, [_D, _D, _D, _P, _P, _P, _P, _I]),
    'gsmi_exponential_cliprows': (_I, [_D, _P, _P, _I, _D, _P]),
    'gsmi_kclip_rows': (_I, [_D, _D, _P, _I, _D, _P]),
    'gsmi_spergel_info': (_I, [_D, _P, _P]),
    'gsmi_inclined_exp_info': (_I, [_D, _D, _P, _P]),
    'gsmi_inclined_sersic_info': (_I, [_D, _D, _D, _D, _P, _P]),
    'gsmi_inclined_exp_kvalue': (_I, [_D, _D, _D, _D, _P, _P, _P, _I]),
    'gsmi_spergel_flux_radius': (_I, [_D, _D, _P]),
    'gsmi_spergel_integrated_flux': (_I, [_D, _D, _P]),
    'gsmi_sersic_kvalue': (_I, [_D, _D, _P, _P, _P, _I]),
    'gsmi_sine_integral': (_I, [_P, _P, _I]),
    'gsmi_lanczos_info': (_I, [_I, _I, _P, _P, _P]),
    'gsmi_lanczos_utable': (_I, [_I, _I, _P, _P, _P, _P, _I]),
    'gsmi_draw_fft_f64': (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_draw_fft_f32': (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_fold_colfft_f64': (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _P]),
    'gsmi_fold_colfft_f32': (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _P]),
    'gsmi_row_c2r_f64': (_I, [_P, _I, _I, _I, _P, _P, _P]),
    'gsmi_row_c2r_f32': (_I, [_P, _I, _I, _I, _P, _P, _P]),
    'gsmi_draw_kimage_f64': (_I, [_P, _I, _P, _P, _P, _I, _I, _I, _I, _I, _D, _P, _P]),
    'gsmi_irfft_f64': (_I, [_P, _I, _P, _P, _P]),
    'gsmi_irfft_f64_noalloc': (_I, [_P, _I, _P, _P]),
    'gsmi_apply_kimage_phases_f64': (_I, [_P, _I, _I, _I, _I, _D, _P, _D, _D, _D, _P]),
    'gsmi_wrap_hermx_f64': (_I, [_P, _I, _I, _P, _P]),
    'gsmi_cfft2_f64': (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P]),
    'gsmi_rfft2_f64': (_I, [_P, _I, _I, _I, _I, _P, _P, _P]),
    'gsmi_irfft2_f64': (_I, [_P, _I, _I, _I, _I, _P, _P, _P]),
    'gsmi_hermitian_fill_f64': (_I, [_P, _I, _I, _I, _P, _P]),
    'gsmi_fft_column_passes': (_I, [_I]),
    'gsmi_draw_list_f64': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_draw_list_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_fold_list_f64': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_fold_list_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_rows_list_f64': (_I, [_P, _I, _P, _P, _P, _P]),
    'gsmi_rows_list_f32': (_I, [_P, _I, _P, _P, _P, _P]),
    'gsmi_adjoint_partials': (_I, [_I, _I, _P]),
    'gsmi_adjoint_f64': (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_adjoint_f32': (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'gsmi_exp_moffat_fields': (_I, [_P, _I, _I, _P, _P, _D, _P, _P, _P, _P, _P]),
    'gsmi_exp_moffat_grad': (_I, [_P, _I, _I, _P, _P, _P, _D, _P, _P]),
    'gsmi_draw_real_f64': (_I, [_P, _I, _P, _I, _I, _I, _P, _P]),
    'gsmi_kolmogorov_xtable': (_I, [_P, _P, _P, _P, _I]),
}


def lib():
    """The loaded library; raises GalSimAMDError if it has not been built."""
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime (SONAME libamdhip64.so.7).  Loading torch first makes
        # the dynamic loader resolve libgsmi's libamdhip64.so.7 to that same copy, so device
        # pointers and streams are shared with torch; a second runtime in the process would not
        # see torch's allocations.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise GalSimAMDError(
                "galsim_amd native library not found at %s; build it with "
                "`python -m galsim_amd.csrc.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().gsmi_last_error().decode(errors='replace')
        text = "%s failed (%s): %s" % (what, ERRORS.get(rc, rc), msg)
        if rc == -2:
            raise GalSimAMDRangeError(text)
        raise GalSimAMDError(text)


def ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def gsp_array(gsparams):
    if gsparams is None:
        return None
    return gsparams._f64_array()
