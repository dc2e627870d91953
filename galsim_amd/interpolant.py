"""Interpolants of InterpolatedImage (galsim/interpolant.py; src/Interpolant.cpp,
include/galsim/Interpolant.h): Nearest, Linear, Cubic, Quintic and Lanczos(n, conserve_dc).

Each carries the code and constants the GPU k-space fill reads (its x-space kernel xval as the
k-interpolant, its Fourier transform uval as the x-interpolant).  ``xval`` / ``kval`` here are
host evaluations for the Python API (the GPU kernels evaluate their own)."""
import math

import numpy as np

from . import _lib
from .gsparams import GSParams


def _sinc(x):
    """math::sinc (src/math/Sinc.cpp:29-33)."""
    x = np.asarray(x, dtype=float)
    with np.errstate(invalid='ignore', divide='ignore'):
        big = np.sin(np.pi * x) / (np.pi * x)
    return np.where(np.abs(x) < 1.e-4, 1. - (np.pi * np.pi / 6.) * x * x, big)


class Interpolant:
    """Base class (galsim/interpolant.py:24-240)."""
    _code = None

    def __init__(self, gsparams=None):
        self._gsparams = GSParams.check(gsparams)

    @property
    def gsparams(self):
        return self._gsparams

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        ret._reset()
        return ret

    def _reset(self):
        pass

    @staticmethod
    def from_name(name, tol=None, gsparams=None):
        """Interpolant.from_name (galsim/interpolant.py:45-110): 'nearest', 'linear', 'cubic',
        'quintic', 'lanczosN' (optionally with a trailing T/F for conserve_dc, default True).
        'delta' and 'sinc' are not part of the GPU path."""
        if tol is not None:
            gsparams = GSParams(kvalue_accuracy=tol)
        gsparams = GSParams.check(gsparams)
        low = name.lower()
        if low == 'quintic':
            return Quintic(gsparams=gsparams)
        if low.startswith('lanczos'):
            conserve_dc = True
            if name[-1].upper() in ('T', 'F'):
                conserve_dc = name[-1].upper() == 'T'
                name = name[:-1]
            try:
                n = int(name[7:])
            except ValueError:
                raise ValueError("Invalid Lanczos specification %s.  Should look like "
                                 "lanczosN, where N is an integer" % name) from None
            return Lanczos(n, conserve_dc, gsparams=gsparams)
        if low == 'linear':
            return Linear(gsparams=gsparams)
        if low == 'cubic':
            return Cubic(gsparams=gsparams)
        if low == 'nearest':
            return Nearest(gsparams=gsparams)
        if low in ('delta', 'sinc'):
            raise NotImplementedError("the %s interpolant is not part of galsim_amd's "
                                      "InterpolatedImage" % low)
        raise ValueError("Invalid Interpolant name %s." % name)

    @property
    def krange(self):
        """2 pi urange: the k beyond which the transform is taken as 0."""
        return 2. * math.pi * self.urange

    def kval(self, k):
        """The Fourier transform at k (uval at k / 2 pi)."""
        return self.uval(np.asarray(k, dtype=float) / (2. * math.pi))

    def _device(self):
        """(code, n, conserve_dc, K(1..5), umax, utable or None) for the parameter block."""
        return self._code, 0, False, (0.,) * 5, 0., None

    def __eq__(self, other):
        return type(self) is type(other) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())

    def _key(self):
        return (type(self).__name__, self.gsparams)


class Nearest(Interpolant):
    """src/Interpolant.cpp:171-176; include/galsim/Interpolant.h:265-290."""
    _code = 0
    xrange = 0.5
    ixrange = 1

    @property
    def urange(self):
        return 1. / (math.pi * self.gsparams.kvalue_accuracy)

    def xval(self, x):
        return np.where(np.abs(x) > 0.5, 0., 1.)

    def uval(self, u):
        return _sinc(u)


class Linear(Interpolant):
    """src/Interpolant.cpp:245-255; include/galsim/Interpolant.h:334-360."""
    _code = 1
    xrange = 1.
    ixrange = 2

    @property
    def urange(self):
        return math.sqrt(1. / self.gsparams.kvalue_accuracy) / math.pi

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        return np.where(x > 1., 0., 1. - x)

    def uval(self, u):
        s = _sinc(u)
        return s * s


class Cubic(Interpolant):
    """src/Interpolant.cpp:284-370 (closed-form transform; uMax where |ft| <= kvalue_accuracy)."""
    _code = 2
    xrange = 2.
    ixrange = 4

    @property
    def urange(self):
        return (3. * math.sqrt(3.) / 8. / self.gsparams.kvalue_accuracy) ** (1. / 3.) / math.pi

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        return np.where(x < 1., 1. + x * x * (1.5 * x - 2.5),
                        np.where(x < 2., -0.5 * (x - 1.) * (x - 2.) * (x - 2.), 0.))

    def uval(self, u):
        u = np.abs(np.asarray(u, dtype=float))
        s = _sinc(u)
        return s * s * s * (3. * s - 2. * np.cos(np.pi * u))


class Quintic(Interpolant):
    """src/Interpolant.cpp:389-531 (Bernstein & Gruen's quintic; closed-form transform)."""
    _code = 3
    xrange = 3.
    ixrange = 6

    @property
    def urange(self):
        return (25. * math.sqrt(5.) / 108. / self.gsparams.kvalue_accuracy) ** (1. / 3.) / math.pi

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        r1 = 1. + x * x * x * (-95. / 12. + x * (23. / 2. + x * (-55. / 12.)))
        r2 = (x - 1.) * (x - 2.) * (-23. / 4. + x * (29. / 2. + x * (-83. / 8. + x * (55. / 24.))))
        r3 = (x - 2.) * (x - 3.) * (x - 3.) * (-9. / 4. + x * (25. / 12. + x * (-11. / 24.)))
        return np.where(x <= 1., r1, np.where(x <= 2., r2, np.where(x <= 3., r3, 0.)))

    def uval(self, u):
        u = np.abs(np.asarray(u, dtype=float))
        s = _sinc(u)
        piu = np.pi * u
        ssq = s * s
        piusq = piu * piu
        return s * ssq * ssq * (s * (55. - 19. * piusq) + 2. * np.cos(piu) * (piusq - 27.))


class Lanczos(Interpolant):
    """Lanczos(n, conserve_dc) (src/Interpolant.cpp:563-919): xval = n sin(pi x) sin(pi x/n) /
    (pi x)^2 (divided by the flux-conserving factor of K(1..5) with conserve_dc); uval from the
    spline table of its transform that the host library builds as the reference does."""
    _code = 4

    def __init__(self, n, conserve_dc=True, tol=None, gsparams=None):
        if tol is not None:
            gsparams = GSParams(kvalue_accuracy=tol)
        super().__init__(gsparams)
        n = int(n)
        if n < 1:
            raise ValueError("Lanczos order must be >= 1")
        if 2 * n + 1 > _lib.INTERP_MAX_TAPS:
            raise NotImplementedError("Lanczos order %d: the GPU k-space fill takes orders up to "
                                      "%d" % (n, (_lib.INTERP_MAX_TAPS - 1) // 2))
        self._n = n
        self._conserve_dc = bool(conserve_dc)
        self._reset()

    def _reset(self):
        self._info = None

    @property
    def n(self):
        return self._n

    @property
    def conserve_dc(self):
        return self._conserve_dc

    @property
    def xrange(self):
        return float(self._n)

    @property
    def ixrange(self):
        return 2 * self._n

    def _table(self):
        if self._info is None:
            lib = _lib.lib()
            out = np.zeros(6)
            ntab = np.zeros(1, dtype=np.int32)
            gsp = _lib.gsp_array(self.gsparams)
            _lib.check(lib.gsmi_lanczos_info(self._n, int(self._conserve_dc), _lib.ptr(gsp),
                                             _lib.ptr(out), _lib.ptr(ntab)), "Lanczos")
            m = int(ntab[0])
            args, vals, y2 = np.zeros(m), np.zeros(m), np.zeros(m)
            _lib.check(lib.gsmi_lanczos_utable(self._n, int(self._conserve_dc), _lib.ptr(gsp),
                                               _lib.ptr(args), _lib.ptr(vals), _lib.ptr(y2), m),
                       "Lanczos table")
            self._info = (tuple(out[:5]), float(out[5]), np.concatenate([args, vals, y2]))
        return self._info

    @property
    def urange(self):
        return self._table()[1]

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        nd = float(self._n)
        with np.errstate(invalid='ignore', divide='ignore'):
            s = np.sin(np.pi * x)
            res = (nd / (np.pi * np.pi)) * s * np.sin(np.pi * x / nd) / (x * x)
        pix = np.pi * x
        temp = (1. / 6.) * pix * pix
        small = x <= 1.e-4
        s = np.where(small, pix * (1. - temp), s)
        res = np.where(small, 1. - temp * (1. + 1. / (nd * nd)), res)
        if self._conserve_dc:
            K = (0.,) + self._table()[0]
            ssq = s * s
            res = res / (1. - 4. * K[1] * ssq - 16. * K[2] * ssq * (1. - ssq)
                         - 4. * K[3] * ssq * (9. - ssq * (24. - 16. * ssq))
                         - 64. * K[4] * ssq * (1. - ssq * (5. - ssq * (8. - 4. * ssq)))
                         - 4. * K[5] * ssq * (25. - ssq * (200. - ssq * (560. - ssq * (
                             640. - 256. * ssq)))))
        return np.where(x >= nd, 0., res)

    def uval(self, u):
        u = np.abs(np.asarray(u, dtype=float))
        _, umax, tab = self._table()
        m = len(tab) // 3
        args, vals, y2 = tab[:m], tab[m:2 * m], tab[2 * m:]
        da = (args[-1] - args[0]) / (m - 1)
        i = np.clip(np.ceil((u - args[0]) / da).astype(np.int64), 1, m - 1)
        h = args[i] - args[i - 1]
        aa = args[i] - u
        bb = h - aa
        v = (aa * vals[i - 1] + bb * vals[i] -
             (1. / 6.) * aa * bb * ((aa + h) * y2[i - 1] + (bb + h) * y2[i])) / h
        return np.where(u > umax, 0., v)

    def _device(self):
        K, umax, tab = self._table()
        return self._code, self._n, self._conserve_dc, K, umax, tab

    def _key(self):
        return ('Lanczos', self._n, self._conserve_dc, self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Lanczos(%r, %r)' % (self._n, self._conserve_dc)


def convert_interpolant(interpolant):
    """An Interpolant as given, or one made from its name (galsim/interpolant.py:672-685)."""
    if isinstance(interpolant, Interpolant):
        return interpolant
    return Interpolant.from_name(interpolant)
