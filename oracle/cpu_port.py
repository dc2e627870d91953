"""ctypes driver of oracle/cpu_draw.cpp: the bulge + disk workload drawn by the C++ restatement
of the oracle on the host's cores (the CPU baseline), with the per-stamp planning (sizes,
leaf constants, Sersic tables) taken from the numpy oracle.

Test infrastructure only (see ``oracle/__init__.py``)."""
import ctypes
import os
import subprocess

import numpy as np

from . import draw as od

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'cpu_draw.cpp')
LIB = os.path.join(HERE, 'libcpudraw.so')

_D = ctypes.c_double
_I = ctypes.c_int32


class BDStamp(ctypes.Structure):
    _fields_ = [('N', _I), ('Nk', _I), ('nx', _I), ('ny', _I), ('xmin', _I), ('ymin', _I),
                ('ser_ntab', _I), ('pad', _I), ('dk', _D), ('inv_scale', _D), ('ox', _D),
                ('oy', _D), ('fs_outer', _D), ('shear', _D * 4), ('fs_shear', _D),
                ('exp_r0', _D), ('exp_flux', _D), ('exp_ksq_max', _D), ('exp_k_max', _D),
                ('ser_r0', _D), ('ser_flux', _D), ('ser_ksq_min', _D), ('ser_ksq_max', _D),
                ('ser_highk_a', _D), ('ser_highk_b', _D), ('ser_kderiv2', _D),
                ('ser_kderiv4', _D), ('kol_k0', _D), ('kol_flux', _D), ('kvalue_accuracy', _D),
                ('ser_table', ctypes.c_void_p), ('out', ctypes.c_void_p)]


def build(force=False):
    """Compile cpu_draw.cpp into oracle/libcpudraw.so (g++, OpenMP, no FP contraction)."""
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    subprocess.run(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-fopenmp',
                    '-ffp-contract=off', SRC, '-o', LIB], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_bd_draw.restype = ctypes.c_int
        L.oracle_bd_draw.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


class Batch:
    """Planned stamps of the bulge + disk workload: ``profiles`` are oracle Convolve([Sum(
    Exponential, Sersic).shear(...), Kolmogorov]) objects, drawn on nx x ny at ``scale`` with
    ``offsets`` by drawImage(method='fft') (oracle/draw.py:194-258)."""

    def __init__(self, profiles, nx, ny, scale, offsets):
        n = len(profiles)
        self.recs = (BDStamp * n)()
        self.outs = []
        self._keep = []
        for i, prof in enumerate(profiles):
            self._plan(self.recs[i], prof, int(nx[i]), int(ny[i]), float(scale),
                       offsets[i])
        self.n = n

    def _plan(self, r, prof, nx, ny, scale, offset):
        gal, kol = prof.obj_list
        summ = gal.original
        exp, ser = summ.obj_list[0].prof, summ.obj_list[1].prof
        kolp = kol.prof
        ox, oy = float(offset[0]), float(offset[1])
        # _adjust_offset with use_true_center
        if nx % 2 == 0:
            ox -= 0.5
        if ny % 2 == 0:
            oy -= 0.5
        j = None if scale == 1. else np.array([[1. / scale, 0.], [0., 1. / scale]])
        T = od._Transform(prof, j, (ox, oy), scale ** 2)
        top = od.Convolution([T, od.Pixel(1.0, gsparams=prof.gsparams)],
                             gsparams=prof.gsparams)
        cx, cy = 1 + nx // 2, 1 + ny // 2
        xmin, xmax, ymin, ymax = 1 - cx, nx - cx, 1 - cy, ny - cy
        # drawFFT's sizes (oracle/draw.py:236-249)
        N = top.get_good_image_size(1.0)
        image_N = max(max(abs(v) for v in (xmin, xmax, ymin, ymax)) * 2, ymax - ymin + 1,
                      xmax - xmin + 1)
        N = max(N, image_N)
        from .numerics import good_fft_size
        N = good_fft_size(N)
        N = max(N, top.gsparams.minimum_fft_size)
        dk = 2. * np.pi / (N * 1.0)
        maxk = top.maxk
        Nk = N if N * dk / 2 > maxk else int(np.ceil(maxk / dk)) * 2
        r.N, r.Nk, r.nx, r.ny, r.xmin, r.ymin = N, Nk, nx, ny, xmin, ymin
        r.dk = dk
        r.inv_scale = 1. / scale
        r.ox, r.oy = ox, oy
        r.fs_outer = T._flux_scaling
        S = gal.jac
        r.shear = (_D * 4)(S[0, 0], S[0, 1], S[1, 0], S[1, 1])
        r.fs_shear = gal._flux_scaling
        r.exp_r0, r.exp_flux, r.exp_ksq_max, r.exp_k_max = exp.r0, exp.flux, exp.ksq_max, \
            exp.k_max
        info = ser.info
        tab = np.ascontiguousarray(np.concatenate([info.table.args, info.table.vals,
                                                   info.table.y2]))
        self._keep.append(tab)
        r.ser_r0, r.ser_flux = ser.r0, ser.flux
        r.ser_ksq_min, r.ser_ksq_max = info.ksq_min, info.ksq_max
        r.ser_highk_a, r.ser_highk_b = info.highk_a, info.highk_b
        r.ser_kderiv2, r.ser_kderiv4 = info.kderiv2, info.kderiv4
        r.ser_ntab = len(info.table.args)
        r.ser_table = tab.ctypes.data
        r.kol_k0, r.kol_flux = kolp.k0, kolp.flux
        r.kvalue_accuracy = prof.gsparams.kvalue_accuracy
        out = np.zeros((ny, nx))
        self.outs.append(out)
        r.out = out.ctypes.data

    def draw(self, nthreads=1, first=0, count=None):
        """Draw stamps [first, first + count) on ``nthreads`` threads (one stamp per thread)."""
        count = self.n - first if count is None else count
        base = ctypes.addressof(self.recs) + first * ctypes.sizeof(BDStamp)
        rc = lib().oracle_bd_draw(base, int(count), int(nthreads))
        if rc != 0:
            raise RuntimeError("oracle_bd_draw failed")
        return self.outs[first:first + count]
