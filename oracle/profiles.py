"""Oracle leaf profiles: size scalars and k-space fills.

Test infrastructure only (see ``oracle/__init__.py``).

Each leaf exposes ``flux``, ``maxk``, ``stepk`` (in the leaf's own length unit) and two fills
that restate the reference's ``fillKImage`` overloads:

* ``fill_1d(kx0, dkx, m, ky0, dky, n)`` -- the separable/diagonal fill (no quadrant logic),
* ``fill_2d(kx0, dkx, dkxy, ky0, dky, dkyx, m, n)`` -- the general-Jacobian fill.

Coordinates are accumulated with ``np.cumsum`` which adds sequentially, i.e. with the same
rounding as the reference's ``kx += dkx`` loops.
"""
import math
import warnings
from collections import deque

import numpy as np
import scipy.special as sps

from . import numerics as nm


class GSParams:
    """galsim/gsparams.py:126-158 defaults."""

    def __init__(self, minimum_fft_size=128, maximum_fft_size=8192, folding_threshold=5.e-3,
                 stepk_minimum_hlr=5, maxk_threshold=1.e-3, kvalue_accuracy=1.e-5,
                 xvalue_accuracy=1.e-5, table_spacing=1, realspace_relerr=1.e-4,
                 realspace_abserr=1.e-6, integration_relerr=1.e-6, integration_abserr=1.e-8,
                 shoot_accuracy=1.e-5):
        self.minimum_fft_size = int(minimum_fft_size)
        self.maximum_fft_size = int(maximum_fft_size)
        self.folding_threshold = float(folding_threshold)
        self.stepk_minimum_hlr = float(stepk_minimum_hlr)
        self.maxk_threshold = float(maxk_threshold)
        self.kvalue_accuracy = float(kvalue_accuracy)
        self.xvalue_accuracy = float(xvalue_accuracy)
        self.table_spacing = int(table_spacing)
        self.realspace_relerr = float(realspace_relerr)
        self.realspace_abserr = float(realspace_abserr)
        self.integration_relerr = float(integration_relerr)
        self.integration_abserr = float(integration_abserr)
        self.shoot_accuracy = float(shoot_accuracy)

    def key(self):
        return tuple(sorted(self.__dict__.items()))

    def __eq__(self, other):
        return isinstance(other, GSParams) and self.key() == other.key()

    def __hash__(self):
        return hash(self.key())

    @staticmethod
    def combine(lst):
        """galsim/gsparams.py:217-241."""
        if len(lst) == 1 or all(g == lst[0] for g in lst[1:]):
            return lst[0]
        return GSParams(
            max(g.minimum_fft_size for g in lst), max(g.maximum_fft_size for g in lst),
            min(g.folding_threshold for g in lst), max(g.stepk_minimum_hlr for g in lst),
            min(g.maxk_threshold for g in lst), min(g.kvalue_accuracy for g in lst),
            min(g.xvalue_accuracy for g in lst), min(g.table_spacing for g in lst),
            min(g.realspace_relerr for g in lst), min(g.realspace_abserr for g in lst),
            min(g.integration_relerr for g in lst), min(g.integration_abserr for g in lst),
            min(g.shoot_accuracy for g in lst))


DEFAULT_GSPARAMS = GSParams()


def _axis(x0, dx, m):
    """x0, x0+dx, ... accumulated left to right (the reference's ``x += dx``)."""
    if m <= 0:
        return np.zeros(0)
    seq = np.empty(m)
    seq[0] = x0
    seq[1:] = dx
    return np.cumsum(seq)


def _grid2d(kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
    """Row starts accumulate by (dkxy, dky) per row, then (dkx, dkyx) per column."""
    rx = _axis(kx0, dkxy, n)
    ry = _axis(ky0, dky, n)
    kx = np.empty((n, m))
    ky = np.empty((n, m))
    kx[:, 0] = rx
    ky[:, 0] = ry
    kx[:, 1:] = dkx
    ky[:, 1:] = dkyx
    return np.cumsum(kx, axis=1), np.cumsum(ky, axis=1)


def fast_pow(x, y):
    """``fmath::expd(y*log(x))`` (src/SBMoffat.cpp:46-47, src/SBKolmogorov.cpp:48-49,
    src/SBSersic.cpp:33-34)."""
    with np.errstate(divide='ignore', invalid='ignore'):
        return nm.expd(y * np.log(x))


def sinc(x):
    """src/math/Sinc.cpp:29-33."""
    x = np.asarray(x, dtype=float)
    with np.errstate(invalid='ignore', divide='ignore'):
        big = np.sin(np.pi * x) / (np.pi * x)
    return np.where(np.abs(x) < 1.e-4, 1. - (np.pi * np.pi / 6.) * x * x, big)


# ----------------------------------------------------------------------------------------------
# Exponential -- src/SBExponential.cpp, galsim/exponential.py
# ----------------------------------------------------------------------------------------------
def _kvalue_range_1d(m, kmax, ksqmax, kx0, dkx, ky):
    """GetKValueRange1d (src/SBProfile.cpp:508-531). Returns (i1, i2, kysq)."""
    if abs(ky) >= kmax:
        return m, m, 0.
    kysq = ky * ky
    dsq = ksqmax - kysq
    if kx0 * kx0 > dsq or (kx0 + m * dkx) * (kx0 + m * dkx) > dsq:
        d = math.sqrt(dsq)
        i1 = int(math.ceil((-kx0 - d) / dkx))
        i2 = int(math.floor((-kx0 + d) / dkx))
        if i1 > i2:
            i1, i2 = i2, i1
        i2 += 1
        if i2 <= 0 or i1 >= m:
            return m, m, kysq
        return max(i1, 0), min(i2, m), kysq
    return 0, m, kysq


def _kvalue_range_2d(m, kmax, ksqmax, kx0, dkx, ky0, dky):
    """GetKValueRange2d (src/SBProfile.cpp:533-588), including the ``ky0+m+dky`` test of 544."""
    if (abs(kx0) > kmax or abs(kx0 + m * dkx) > kmax or
            abs(ky0) > kmax or abs(ky0 + m + dky) > kmax):
        a = dkx * dkx + dky * dky
        b = dkx * kx0 + dky * ky0
        c = kx0 * kx0 + ky0 * ky0 - ksqmax
        d = b * b - a * c
        if d <= 0.:
            return m, m
        d = math.sqrt(d)
        i1 = int(math.ceil((-b - d) / a))
        i2 = int(math.floor((-b + d) / a))
        if i1 > i2:
            i1, i2 = i2, i1
        i2 += 1
        if i2 <= 0 or i1 >= m:
            return m, m
        return max(i1, 0), min(i2, m)
    return 0, m


class Exponential:
    """SBExponential (src/SBExponential.cpp:64-92 ctor, 442-510 fills, 513-545 info).

    Pinned by the closed form Exponential(k) == Moffat(beta=1.5).xValue (tests/test_draw.py:859)
    and stepk*r0 = 0.37436747851, maxk*r0 = 10 (tests/test_exponential.py:100-101).
    """
    hlr_factor = 1.6783469900166605

    def __init__(self, half_light_radius=None, scale_radius=None, flux=1., gsparams=None):
        if half_light_radius is not None:
            scale_radius = half_light_radius / Exponential.hlr_factor
        self.r0 = float(scale_radius)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        kva = self.gsparams.kvalue_accuracy
        self.ksq_max = kva ** (-1. / 1.5) - 1.
        self.k_max = math.sqrt(self.ksq_max)

    @property
    def maxk(self):
        # galsim/exponential.py:131-132
        return (self.gsparams.maxk_threshold ** (-1. / 3.)) / self.r0

    @property
    def stepk(self):
        # ExponentialInfo ctor, src/SBExponential.cpp:527-544
        gsp = self.gsparams
        logx = math.log(gsp.folding_threshold)
        R = -logx
        for _ in range(3):
            R = math.log(1. + R) - logx
        R = max(R, gsp.stepk_minimum_hlr * 1.6783469900166605)
        return (math.pi / R) / self.r0

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; ky0 *= r0; dky *= r0
        out = np.zeros((n, m))
        kys = _axis(ky0, dky, n)
        for j in range(n):
            i1, i2, kysq = _kvalue_range_1d(m, self.k_max, self.ksq_max, kx0, dkx, kys[j])
            if i1 == m:
                continue
            kx = _axis(kx0 + i1 * dkx, dkx, i2 - i1)
            ksqp1 = kx * kx + (kysq + 1.)
            out[j, i1:i2] = self.flux / (ksqp1 * np.sqrt(ksqp1))
        return out

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; dkxy *= r0; ky0 *= r0; dky *= r0; dkyx *= r0
        out = np.zeros((n, m))
        rx = _axis(kx0, dkxy, n)
        ry = _axis(ky0, dky, n)
        for j in range(n):
            i1, i2 = _kvalue_range_2d(m, self.k_max, self.ksq_max, rx[j], dkx, ry[j], dkyx)
            if i1 == m:
                continue
            kx = _axis(rx[j] + i1 * dkx, dkx, i2 - i1)
            ky = _axis(ry[j] + i1 * dkyx, dkyx, i2 - i1)
            ksqp1 = 1. + kx * kx + ky * ky
            out[j, i1:i2] = self.flux / (ksqp1 * np.sqrt(ksqp1))
        return out


# ----------------------------------------------------------------------------------------------
# Spergel -- src/SBSpergel.cpp, galsim/spergel.py
# ----------------------------------------------------------------------------------------------
def spergel_integrated_flux(nu, u):
    """Enclosed flux fraction within u r0 (SpergelIntegratedFlux, src/SBSpergel.cpp:545-580;
    Spergel 2010 eq. 8): 1 - 2 (1+nu) (u/2)^(nu+1) K_{nu+1}(u) / Gamma(nu+2)."""
    return 1. - 2. * (1. + nu) * (u / 2.) ** (nu + 1.) * sps.kv(nu + 1., u) / math.gamma(nu + 2.)


def spergel_flux_radius(nu, f):
    """CalculateFluxRadius (src/SBSpergel.cpp:582-598), solved to machine precision (the
    reference asks its Brent solver for an x tolerance of 1e-25)."""
    import scipy.optimize as spo
    func = lambda u: spergel_integrated_flux(nu, u) - f
    lo, hi = 0.1, 3.5
    while func(lo) > 0.:
        lo *= 0.5
    while func(hi) < 0.:
        hi *= 2.
    return spo.brentq(func, lo, hi, xtol=1e-300, rtol=4 * np.finfo(float).eps, maxiter=500)


class Spergel:
    """SBSpergel (src/SBSpergel.cpp:72-99 ctor, 456-527 fills with kloop_1d/2d 121-139,
    529-653 SpergelInfo); galsim/spergel.py:28-239.

    Pinned by the known answers of tests/test_spergel.py:106-190 (scale radius, maxk, stepk at
    nu = 0.2 and -0.2), the Mathematica flux radii of tests/test_spergel.py:30-83, and
    nu = 0.5 being the Exponential (k-image and half-light radius).
    """

    def __init__(self, nu, half_light_radius=None, scale_radius=None, flux=1., gsparams=None):
        self.nu = float(nu)
        if not -0.85 <= self.nu <= 4.0:
            raise ValueError("Spergel index out of range")
        if half_light_radius is not None:
            scale_radius = half_light_radius / spergel_flux_radius(self.nu, 0.5)
        self.r0 = float(scale_radius)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        self.ksq_max = self.gsparams.kvalue_accuracy ** (-1. / (self.nu + 1.)) - 1.
        self.k_max = math.sqrt(self.ksq_max)

    @property
    def maxk(self):
        # SpergelInfo::maxK (src/SBSpergel.cpp:629-636)
        thr = self.gsparams.maxk_threshold
        return math.sqrt(thr ** (-1. / (1 + self.nu)) - 1.0) / self.r0

    @property
    def stepk(self):
        # SpergelInfo::stepK (src/SBSpergel.cpp:616-627)
        gsp = self.gsparams
        R = spergel_flux_radius(self.nu, 1. - gsp.folding_threshold)
        R = max(R, gsp.stepk_minimum_hlr * spergel_flux_radius(self.nu, 0.5))
        return (math.pi / R) / self.r0

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; ky0 *= r0; dky *= r0
        out = np.zeros((n, m))
        kys = _axis(ky0, dky, n)
        for j in range(n):
            i1, i2, kysq = _kvalue_range_1d(m, self.k_max, self.ksq_max, kx0, dkx, kys[j])
            if i1 == m:
                continue
            kx = _axis(kx0 + i1 * dkx, dkx, i2 - i1)
            ksqp1 = kx * kx + (kysq + 1.)
            out[j, i1:i2] = self.flux * ksqp1 ** (-(self.nu + 1.))
        return out

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; dkxy *= r0; ky0 *= r0; dky *= r0; dkyx *= r0
        out = np.zeros((n, m))
        rx = _axis(kx0, dkxy, n)
        ry = _axis(ky0, dky, n)
        for j in range(n):
            i1, i2 = _kvalue_range_2d(m, self.k_max, self.ksq_max, rx[j], dkx, ry[j], dkyx)
            if i1 == m:
                continue
            kx = _axis(rx[j] + i1 * dkx, dkx, i2 - i1)
            ky = _axis(ry[j] + i1 * dkyx, dkyx, i2 - i1)
            ksqp1 = 1. + kx * kx + ky * ky
            out[j, i1:i2] = self.flux * ksqp1 ** (-(self.nu + 1.))
        return out


# ----------------------------------------------------------------------------------------------
# Airy -- src/SBAiry.cpp, galsim/airy.py
# ----------------------------------------------------------------------------------------------
def _airy_chord(r, h, rsq, hsq):
    """AiryInfoObs::chord (src/SBAiry.cpp:259-272)."""
    if r == 0.:
        return np.zeros_like(h)
    return rsq * np.arcsin(h / r) - h * np.sqrt(rsq - hsq)


def _airy_circles(r, s, rsq, ssq, tsq):
    """Area of the intersection of circles of radii r >= s, centres sqrt(tsq) apart
    (src/SBAiry.cpp:274-296)."""
    rps_sq = (r + s) * (r + s)
    rms_sq = (r - s) * (r - s)
    out = np.zeros_like(tsq)
    inner = tsq <= rms_sq
    out[inner] = math.pi * ssq
    mid = (tsq < rps_sq) & ~inner
    t = tsq[mid]
    hsq = 0.5 * (rsq + ssq) - (t * t + rps_sq * rms_sq) / (4. * t)
    h = np.sqrt(hsq)
    a = _airy_chord(s, h, ssq, hsq) + _airy_chord(r, h, rsq, hsq)
    b = math.pi * ssq - _airy_chord(s, h, ssq, hsq) + _airy_chord(r, h, rsq, hsq)
    out[mid] = np.where(t < rsq - ssq, b, a)
    return out


def _airy_circle_pair(r, rsq, tsq):
    """Equal radii (src/SBAiry.cpp:298-313)."""
    out = np.zeros_like(tsq)
    out[tsq == 0.] = math.pi * rsq
    mid = (tsq < 4. * rsq) & (tsq != 0.)
    hsq = rsq - tsq[mid] / 4.
    out[mid] = 2. * _airy_chord(r, np.sqrt(hsq), rsq, hsq)
    return out


def airy_kvalue(t, obs):
    """Overlap area of two annuli (outer radius 1, inner obs) at separation sqrt(t)
    (AiryInfoObs::kValue src/SBAiry.cpp:315-329; AiryInfoNoObs::kValue 409-420)."""
    t = np.asarray(t, dtype=float)
    if obs == 0.:
        out = np.zeros_like(t)
        out[t == 0.] = math.pi
        mid = (t < 4.) & (t != 0.)
        hsq = 1. - t[mid] / 4.
        h = np.sqrt(hsq)
        out[mid] = 2. * (np.arcsin(h) - h * np.sqrt(1. - hsq))
        return out
    obssq = obs * obs
    return (_airy_circle_pair(1., 1., t) - 2. * _airy_circles(1., obs, 1., obssq, t) +
            _airy_circle_pair(obs, obssq, t))


class Airy:
    """SBAiry (src/SBAiry.cpp:66-80 ctor, 189-248 fills, 250-257 sizes, 332-348 and 445-455
    stepK); galsim/airy.py:28-249.

    Pinned by the reference's known maxk / stepk (tests/test_airy.py:226-240), kValue(0) = flux,
    the textbook lens area of two unit circles (obscuration 0) and a brute-force area integral of
    two overlapping annuli (tests/test_airy_cpu.py).
    """

    def __init__(self, lam_over_diam, obscuration=0., flux=1., gsparams=None):
        self.lod = float(lam_over_diam)
        self.obs = float(obscuration)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        self.knorm = self.flux / (math.pi * (1. - self.obs * self.obs))
        self.inv_D_pi = self.lod / math.pi

    @property
    def maxk(self):
        return 2. * math.pi / self.lod

    @property
    def stepk(self):
        gsp = self.gsparams
        R = 1. / (gsp.folding_threshold * 0.5 * math.pi * math.pi * (1. - self.obs))
        R = max(R, gsp.stepk_minimum_hlr * 0.5348321477)
        return (math.pi / R) / self.lod

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        f = self.inv_D_pi
        kx0 *= f; dkx *= f; ky0 *= f; dky *= f
        kx = _axis(kx0, dkx, m)
        ky = _axis(ky0, dky, n)
        t = kx[None, :] * kx[None, :] + (ky * ky)[:, None]
        return self.knorm * airy_kvalue(t, self.obs)

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        f = self.inv_D_pi
        kx0 *= f; dkx *= f; dkxy *= f; ky0 *= f; dky *= f; dkyx *= f
        out = np.zeros((n, m))
        rx = _axis(kx0, dkxy, n)
        ry = _axis(ky0, dky, n)
        for j in range(n):
            kx = _axis(rx[j], dkx, m)
            ky = _axis(ry[j], dkyx, m)
            out[j] = self.knorm * airy_kvalue(kx * kx + ky * ky, self.obs)
        return out


# ----------------------------------------------------------------------------------------------
# InclinedExponential -- src/SBInclinedExponential.cpp, galsim/inclined.py
# ----------------------------------------------------------------------------------------------
def inclined_exp_kvalue(kx, ky, cosi, hs, ksq_min, ksq_max):
    """kValueHelper (src/SBInclinedExponential.cpp:277-336) at (kx, ky) in units of 1/r0."""
    kx = np.asarray(kx, dtype=float)
    ky = np.asarray(ky, dtype=float)
    ky_cosi = ky * cosi
    ksq = kx * kx + ky_cosi * ky_cosi
    temp = 1. + ksq
    base = np.where(ksq < ksq_min, 1. - 1.5 * ksq * (1. - 1.25 * ksq),
                    1. / (temp * np.sqrt(temp)))
    sky = hs * ky
    sky2 = sky * sky
    with np.errstate(invalid='ignore', divide='ignore', over='ignore'):
        conv = np.where(sky2 < ksq_min, 1. - 0.16666666667 * sky2 * (1. - 0.116666666667 * sky2),
                        sky / np.sinh(sky))
    return np.where(ksq > ksq_max, 0., base * conv)


class InclinedExponential:
    """SBInclinedExponential (src/SBInclinedExponential.cpp:64-160 ctor with its maxk (Brent) and
    clipping-radius (bisection) solves, 207-275 fills, 277-336 kValueHelper); galsim/inclined.py.

    Pinned by Lance Miller's independent images (tests/inclined_exponential_images, compared in
    tests/test_inclined.py:101-160), by inclination 0 reducing to the Exponential and by the
    maxk / clipping roots satisfying their equations.
    """

    def __init__(self, inclination, scale_radius, scale_height, flux=1., gsparams=None):
        self.r0 = float(scale_radius)
        self.h0 = float(scale_height)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        gsp = self.gsparams
        self.cosi = abs(math.cos(inclination))
        self.hs = 0.5 * math.pi * self.h0 * abs(math.sin(inclination)) / self.r0
        logx = math.log(gsp.folding_threshold)
        R = -logx
        for _ in range(3):
            R = math.log(1. + R) - logx
        R = max(R, gsp.stepk_minimum_hlr * 1.6783469900166605)
        self._stepk = math.pi / R
        self.ksq_min = (gsp.kvalue_accuracy / (35. / 16. + 31. / 15120. * self.hs)) ** (1. / 3.)
        maxk_min = gsp.maxk_threshold ** (-1. / 3.)
        clipk_min = gsp.kvalue_accuracy ** (-1. / 3.)
        if self.cosi == 1:
            self._maxk = maxk_min
            self.ksq_max = clipk_min * clipk_min
            return
        if self.cosi > 0.01:
            if self.cosi < 0.96:
                maxk_max, clipk_max = maxk_min / self.cosi, clipk_min / self.cosi
            else:
                maxk_max, clipk_max = 1.05 * maxk_min, 1.05 * clipk_min
        else:
            maxk_max, clipk_max = 100 * maxk_min, 100 * clipk_min

        def func(target):
            def f(k):
                a = float(inclined_exp_kvalue(0., k, self.cosi, self.hs, self.ksq_min, math.inf))
                b = float(inclined_exp_kvalue(k, 0., self.cosi, self.hs, self.ksq_min, math.inf))
                return max(a, b) - target
            return f

        f = func(gsp.maxk_threshold)
        s = nm.Solve(f, maxk_min, maxk_max)
        if f(maxk_min) <= 0:
            s.bracket_lower_with_limit(0.)
        else:
            s.bracket_upper()
        self._maxk = s.root() + s.xtol
        f = func(gsp.kvalue_accuracy)
        s = nm.Solve(f, clipk_min, clipk_max)
        if f(clipk_min) <= 0:
            s.bracket_lower_with_limit(0.)
        else:
            s.bracket_upper()
        clipk = s.bisect() + s.xtol
        self.ksq_max = clipk * clipk

    @property
    def maxk(self):
        return self._maxk / self.r0

    @property
    def stepk(self):
        return self._stepk / self.r0

    def _kv(self, kx, ky):
        return self.flux * inclined_exp_kvalue(kx, ky, self.cosi, self.hs, self.ksq_min,
                                               self.ksq_max)

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; ky0 *= r0; dky *= r0
        kx = _axis(kx0, dkx, m)
        ky = _axis(ky0, dky, n)
        return self._kv(kx[None, :], ky[:, None] + 0. * kx[None, :])

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; dkxy *= r0; ky0 *= r0; dky *= r0; dkyx *= r0
        out = np.zeros((n, m))
        rx = _axis(kx0, dkxy, n)
        ry = _axis(ky0, dky, n)
        for j in range(n):
            out[j] = self._kv(_axis(rx[j], dkx, m), _axis(ry[j], dkyx, m))
        return out


# ----------------------------------------------------------------------------------------------
# Gaussian -- src/SBGaussian.cpp, galsim/gaussian.py
# ----------------------------------------------------------------------------------------------
class Gaussian:
    """SBGaussian (src/SBGaussian.cpp:52-79 ctor, 200-290 fills); maxk/stepk galsim/gaussian.py:147-155.

    Pinned end to end by gauss_smallshear_convolve_box.fits (tests/test_convolve.py:220).
    """
    hlr_factor = 1.1774100225154747
    fwhm_factor = 2.3548200450309493

    def __init__(self, sigma=None, half_light_radius=None, fwhm=None, flux=1., gsparams=None):
        if fwhm is not None:
            sigma = fwhm / Gaussian.fwhm_factor
        elif half_light_radius is not None:
            sigma = half_light_radius / Gaussian.hlr_factor
        self.sigma = float(sigma)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        kva = self.gsparams.kvalue_accuracy
        self.ksq_max = -2. * math.log(kva)
        self.ksq_min = (kva * 48.) ** (1. / 3.)

    @property
    def maxk(self):
        return math.sqrt(-2. * math.log(self.gsparams.maxk_threshold)) / self.sigma

    @property
    def stepk(self):
        R = max(math.sqrt(-2. * math.log(self.gsparams.folding_threshold)),
                self.gsparams.stepk_minimum_hlr * Gaussian.hlr_factor)
        return math.pi / (R * self.sigma)

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        s = self.sigma
        kx = _axis(kx0 * s, dkx * s, m)
        ky = _axis(ky0 * s, dky * s, n)
        gx = np.exp(-0.5 * kx * kx)
        gy = np.exp(-0.5 * ky * ky)
        return (self.flux * gx[None, :]) * gy[:, None]

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        s = self.sigma
        kx, ky = _grid2d(kx0 * s, dkx * s, dkxy * s, ky0 * s, dky * s, dkyx * s, m, n)
        ksq = kx * kx + ky * ky
        out = self.flux * np.exp(-0.5 * ksq)
        out = np.where(ksq < self.ksq_min, self.flux * (1. - 0.5 * ksq * (1. - 0.25 * ksq)), out)
        return np.where(ksq > self.ksq_max, 0., out)


# ----------------------------------------------------------------------------------------------
# Moffat -- src/SBMoffat.cpp, galsim/moffat.py
# ----------------------------------------------------------------------------------------------
def _moffat_trunc_table(beta, maxRrD, gsp):
    """SBMoffatImpl::setupFT (src/SBMoffat.cpp:565-626): kV_trunc tabulated on k^2 for
    k = 0, dk, 2 dk, ... (a running sum; dk = table_spacing (kvalue_accuracy/10)^(1/4)) until 5
    successive |values| <= kvalue_accuracy or k >= 50; value = 2 (beta-1) / fluxFactor times
    int_0^maxRrD (1+r^2)^-beta J0(k r) r dr.  The integrals here are scipy's QUADPACK on the
    reference's split points (the first 10 zeros of J0 below k maxRrD, src/math/Hankel.cpp:37-63)
    to 1e-14; the last k with |value| > maxk_threshold is refined linearly towards the next knot.
    Returns (SplineTable on k^2, maxk in units of 1/rD, fluxFactor).  Pinned against the
    reference's own hankel_trunc (oracle/_ref) and the maxk/stepk known answers of
    tests/test_moffat.py:130-160."""
    import scipy.integrate as spi
    key = (beta, maxRrD, gsp.key())
    if key in _MOFFAT_TRUNC:
        return _MOFFAT_TRUNC[key]
    flux_factor = 1. - math.pow(1. + maxRrD * maxRrD, 1. - beta)
    prefactor = 2. * (beta - 1.) / flux_factor
    zeros = sps.jn_zeros(0, 10)

    def hankel(k):
        f = lambda r: r * math.exp(-beta * math.log1p(r * r)) * sps.j0(k * r)
        cuts = [0.] + [z / k for z in zeros if k > 0 and z <= k * maxRrD] + [maxRrD]
        with warnings.catch_warnings():
            warnings.simplefilter('ignore')          # QUADPACK's roundoff notes at 1e-14
            return sum(spi.quad(f, a, b, epsabs=1e-15, epsrel=1e-14, limit=400)[0]
                       for a, b in zip(cuts[:-1], cuts[1:]))

    dk = gsp.table_spacing * math.sqrt(math.sqrt(gsp.kvalue_accuracy / 10.))
    args, vals = [], []
    maxk = 0.
    below = 0
    k = 0.
    while k < 50:
        v = prefactor * hankel(k)
        args.append(k * k)
        vals.append(v)
        if abs(v) > gsp.maxk_threshold:
            maxk = k
        below = 0 if abs(v) > gsp.kvalue_accuracy else below + 1
        if below == 5:
            break
        k += dk
    ft = nm.SplineTable(args, vals)
    k1, k2 = maxk, maxk + dk
    if k2 < 50:
        f1, f2 = abs(float(ft(k1 * k1))), abs(float(ft(k2 * k2)))
        t = gsp.maxk_threshold
        maxk = (k1 * (f2 - t) + k2 * (t - f1)) / (f2 - f1)
    out = (ft, maxk, flux_factor)
    _MOFFAT_TRUNC[key] = out
    return out


_MOFFAT_TRUNC = {}


class Moffat:
    """SBMoffat (src/SBMoffat.cpp:163-244 ctor, 286-345 kV_*, 419-479 fills, 497-549
    maxK/stepK, 565-626 the truncated table).  K_nu from scipy.special.kv.

    Pinned by kValue(maxk)/flux == maxk_threshold (tests/test_moffat.py:183-214), the drawKImage
    size N=1174 for beta=1.5, maxk_threshold=1e-4 (tests/test_draw.py:1279) and the Hankel pair
    with the Exponential (tests/test_draw.py:859); truncated, see _moffat_trunc_table.
    """

    def __init__(self, beta, scale_radius=None, half_light_radius=None, fwhm=None, trunc=0.,
                 flux=1., gsparams=None):
        self.beta = float(beta)
        self.trunc = float(trunc)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        if half_light_radius is not None and self.trunc > 0.:
            # MoffatCalculateScaleRadiusFromHLR with rm = trunc (src/SBMoffat.cpp:127-161)
            re, rm, b = float(half_light_radius), self.trunc, self.beta
            if rm <= math.sqrt(2.) * re:
                raise ValueError("Moffat trunc must be > sqrt(2) * half_light_radius.")

            def func(rd):
                fre = 1. - fast_pow(1. + re * re / (rd * rd), 1. - b)
                frm = 1. - fast_pow(1. + rm * rm / (rd * rd), 1. - b)
                return 2. * fre - frm
            r1 = re / math.sqrt(math.pow(0.5, 1. / (1. - b)) - 1.)
            s = nm.Solve(func, r1, 2. * r1)
            s.bracket_upper()
            scale_radius = s.root()
        elif half_light_radius is not None:
            # MoffatCalculateScaleRadiusFromHLR, rm == 0 branch (src/SBMoffat.cpp:116-126)
            rerd = math.sqrt(math.pow(0.5, 1. / (1. - self.beta)) - 1.)
            scale_radius = half_light_radius / rerd
        elif fwhm is not None:
            # galsim/moffat.py:111
            scale_radius = fwhm / (2. * math.sqrt(2. ** (1. / self.beta) - 1.))
        self.rD = float(scale_radius)
        gsp = self.gsparams
        beta = self.beta
        self._maxk = None
        if self.trunc > 0.:
            self.maxRrD = self.trunc / self.rD
            self.maxR = self.trunc
            self.ft, self._maxk, self.fluxFactor = _moffat_trunc_table(beta, self.maxRrD, gsp)
            self.kind = 1
            self.knorm2 = 1.
            self.knorm = self.flux
            return
        self.maxRrD = math.sqrt(math.pow(gsp.xvalue_accuracy, 1. / (1. - beta)) - 1.)
        self.maxR = self.maxRrD * self.rD
        self.fluxFactor = 1.
        kva = gsp.kvalue_accuracy
        self.knorm2 = 1.
        if abs(beta - 1.5) < kva:
            self.kind = 15
        elif abs(beta - 2) < kva:
            self.kind = 2
        elif abs(beta - 2.5) < kva:
            self.kind = 25
        elif abs(beta - 3) < kva:
            self.kind = 3
            self.knorm2 /= 2.
        elif abs(beta - 3.5) < kva:
            self.kind = 35
            self.knorm2 /= 3.
        elif abs(beta - 4) < kva:
            self.kind = 4
            self.knorm2 /= 8.
        else:
            self.kind = 0
            self.knorm2 *= 4. / (math.gamma(beta - 1.) * math.pow(2., beta))
        self.knorm = self.flux * self.knorm2

    def kV(self, ksq):
        """kV_* of src/SBMoffat.cpp:286-345 (argument ksq = k^2 rD^2)."""
        ksq = np.asarray(ksq, dtype=float)
        k = np.sqrt(ksq)
        kind = self.kind
        if kind == 1:                       # kV_trunc (340-345): 0 beyond the table
            return np.where(ksq > self.ft.arg_max, 0., self.ft(ksq))
        with np.errstate(invalid='ignore', divide='ignore', over='ignore'):
            if kind == 15:
                return nm.expd(-k)
            if kind == 25:
                return nm.expd(-k) * (1. + k)
            if kind == 35:
                return nm.expd(-k) * (3. + (3. + k) * k)
            if kind == 2:
                return np.where(ksq == 0., 1., sps.kv(1, k) * k)
            if kind == 3:
                return np.where(ksq == 0., 2., sps.kv(2, k) * ksq)
            if kind == 4:
                return np.where(ksq == 0., 8., sps.kv(3, k) * k * ksq)
            nu = self.beta - 1.
            return np.where(ksq == 0., self.flux / self.knorm, sps.kv(nu, k) * fast_pow(k, nu))

    @property
    def half_light_radius(self):
        return self.rD * math.sqrt(math.pow(1. - 0.5 * self.fluxFactor,
                                            1. / (1. - self.beta)) - 1.)

    @property
    def maxk(self):
        # SBMoffatImpl::maxK (src/SBMoffat.cpp:497-518): Brent on kV2(ksq) - maxk_threshold.
        if self._maxk is None:
            thresh = self.gsparams.maxk_threshold
            func = lambda ksq: float(self.knorm2 * self.kV(ksq)) - thresh
            s = nm.Solve(func, 0., 100.)
            s.bracket_upper()
            self._maxk = math.sqrt(s.root())
        return self._maxk / self.rD

    @property
    def stepk(self):
        # src/SBMoffat.cpp:522-549 (maxR is the truncation radius when truncated)
        gsp = self.gsparams
        if self.beta <= 1.1:
            return math.pi / self.maxR
        R = math.pow(gsp.folding_threshold, 0.5 / (1. - self.beta)) * self.rD
        if R > self.maxR:
            R = self.maxR
        R = max(R, gsp.stepk_minimum_hlr * self.half_light_radius)
        return math.pi / R

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        rD = self.rD
        kx = _axis(kx0 * rD, dkx * rD, m)
        ky = _axis(ky0 * rD, dky * rD, n)
        ksq = kx[None, :] * kx[None, :] + (ky * ky)[:, None]
        return self.knorm * self.kV(ksq)

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        rD = self.rD
        kx, ky = _grid2d(kx0 * rD, dkx * rD, dkxy * rD, ky0 * rD, dky * rD, dkyx * rD, m, n)
        return self.knorm * self.kV(kx * kx + ky * ky)


# ----------------------------------------------------------------------------------------------
# Kolmogorov -- src/SBKolmogorov.cpp, galsim/kolmogorov.py
# ----------------------------------------------------------------------------------------------
K0_FACTOR = 2.992939911888651
XVAL_ZERO = 0.0876786563672346
_KOLM_INFO = {}


def kolmogorov_info(gsp):
    """KolmogorovInfo ctor (src/SBKolmogorov.cpp:337-419): (maxk, stepk) in units of k0.

    The x-space radial table is the Hankel transform of exp(-k^(5/3)) on log r from -3 in steps
    of table_spacing*(xvalue_accuracy/10)^(1/4), stopped once val r^2/1.67 < shoot_accuracy/2pi;
    stepk follows from the r^(-11/3) tail through the last entry.  Pinned by maxk and stepk at
    lam_over_r0=1.5 (tests/test_kolmogorov.py:138-139) and against oracle/_ref.
    """
    key = gsp.key()
    if key in _KOLM_INFO:
        return _KOLM_INFO[key]
    maxk = math.pow(-math.log(gsp.kvalue_accuracy), 3. / 5.)
    dlogr = gsp.table_spacing * math.sqrt(math.sqrt(gsp.xvalue_accuracy / 10.))
    thresh = gsp.shoot_accuracy / (2. * math.pi)
    kfunc = lambda k: nm.expd(-fast_pow(k, 5. / 3.))
    logr = -3.
    r1 = F1 = None
    while logr < math.log(1.e4):
        r = math.exp(logr)
        val = nm.hankel_inf(kfunc, r, gsp.integration_relerr, gsp.integration_abserr)
        val /= 2. * math.pi
        r1, F1 = r, val
        if val * r * r / 1.67 < thresh:
            break
        logr += dlogr
    n = 11. / 3.
    R = float(fast_pow(2. * math.pi * F1 * float(fast_pow(r1, n)) /
                       ((n - 2) * gsp.folding_threshold), 1. / (n - 2)))
    hlr = K0_FACTOR * 0.5548101137
    R = max(R, gsp.stepk_minimum_hlr * hlr)
    res = (maxk, math.pi / R)
    _KOLM_INFO[key] = res
    return res


def kolmogorov_xtable(gsp):
    """KolmogorovInfo's x-space radial table (src/SBKolmogorov.cpp:337-384): knots (0, then
    exp(log r) from -3 by dlogr until val r^2 / 1.67 < shoot_accuracy / 2 pi) and values (f(0),
    then the Hankel transform of exp(-k^(5/3)) / 2 pi), the table kolmogorov_info walks."""
    key = ('xtable',) + tuple(gsp.key())
    if key in _KOLM_INFO:
        return _KOLM_INFO[key]
    dlogr = gsp.table_spacing * math.sqrt(math.sqrt(gsp.xvalue_accuracy / 10.))
    thresh = gsp.shoot_accuracy / (2. * math.pi)
    kfunc = lambda k: nm.expd(-fast_pow(k, 5. / 3.))
    args, vals = [0.], [0.0876786563672346]
    logr = -3.
    while logr < math.log(1.e4):
        r = math.exp(logr)
        val = nm.hankel_inf(kfunc, r, gsp.integration_relerr, gsp.integration_abserr)
        val /= 2. * math.pi
        args.append(r)
        vals.append(val)
        if val * r * r / 1.67 < thresh:
            break
        logr += dlogr
    res = (np.array(args), np.array(vals))
    _KOLM_INFO[key] = res
    return res


class Kolmogorov:
    """SBKolmogorov (src/SBKolmogorov.cpp:83-100 ctor, 185-245 fills, 258-259 kValue);
    maxk galsim/kolmogorov.py:281-284."""
    fwhm_factor = 0.9758634299
    hlr_factor = 0.5548101137

    def __init__(self, lam_over_r0=None, fwhm=None, half_light_radius=None, flux=1.,
                 gsparams=None):
        if fwhm is not None:
            lam_over_r0 = fwhm / Kolmogorov.fwhm_factor
        elif half_light_radius is not None:
            lam_over_r0 = half_light_radius / Kolmogorov.hlr_factor
        self.lam_over_r0 = float(lam_over_r0)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        self.k0 = K0_FACTOR / self.lam_over_r0

    @property
    def maxk(self):
        return (-math.log(self.gsparams.kvalue_accuracy)) ** 0.6 * self.k0

    @property
    def stepk(self):
        return kolmogorov_info(self.gsparams)[1] * self.k0

    def _kv(self, ksq):
        with np.errstate(divide='ignore'):
            return self.flux * nm.expd(-fast_pow(ksq, 5. / 6.))

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        ik = 1. / self.k0
        kx = _axis(kx0 * ik, dkx * ik, m)
        ky = _axis(ky0 * ik, dky * ik, n)
        return self._kv(kx[None, :] * kx[None, :] + (ky * ky)[:, None])

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        ik = 1. / self.k0
        kx, ky = _grid2d(kx0 * ik, dkx * ik, dkxy * ik, ky0 * ik, dky * ik, dkyx * ik, m, n)
        return self._kv(kx * kx + ky * ky)


# ----------------------------------------------------------------------------------------------
# Sersic -- src/SBSersic.cpp, galsim/sersic.py
# ----------------------------------------------------------------------------------------------
def _gamma_p(a, x):
    return float(sps.gammainc(a, x))


class _SersicMissingFlux:
    """src/SBSersic.cpp:542-556."""

    def __init__(self, n, missing_flux):
        self.two_n = 2. * n
        self.target = missing_flux
        self.g2n = math.gamma(self.two_n)

    def __call__(self, z):
        return (1. - _gamma_p(self.two_n, z)) * self.g2n - self.target


def sersic_calculate_b(n, flux_fraction=1.):
    """CalculateB (src/SBSersic.cpp:627-654)."""
    invn = 1. / n
    gamma2n = math.gamma(2. * n)
    invnsq = invn * invn
    b1 = 2. * n - 1. / 3.
    b2 = b1 + (8. / 405.) * invn + (46. / 25515.) * invnsq + (131. / 1148175.) * invn * invnsq
    func = _SersicMissingFlux(n, (1. - 0.5 * flux_fraction) * gamma2n)
    s = nm.Solve(func, b1, b2)
    s.bracket_lower_with_limit(0.)
    return s.root()


def sersic_hlr(n, flux_fraction=1.):
    """SersicHLR (src/SBSersic.cpp:665-669): re/r0."""
    return math.pow(sersic_calculate_b(n, flux_fraction), n)


def sersic_truncated_scale(n, hlr, trunc):
    """SersicTruncatedScale / CalculateTruncatedScale (src/SBSersic.cpp:693-775)."""
    b = sersic_calculate_b(n, 1.)
    t = trunc / hlr
    if t <= math.sqrt(2.):
        raise ValueError("Sersic truncation must be larger than sqrt(2)*half_light_radius.")
    x = math.pow(t, 1. / n)
    b1 = (math.log(0.5) + (2. * n - 1) * math.log(x)) / (x - 1.)
    if b1 < 1.e-3 * b:
        b1 = b / 2
    two_n = 2. * n

    def func(bb):
        return (2. * _gamma_p(two_n, bb) - _gamma_p(two_n, x * bb)) * math.gamma(two_n)
    s = nm.Solve(func, b1, b)
    s.bracket_lower_with_limit(0.)
    return hlr / math.pow(s.root(), n)


_GL24 = np.polynomial.legendre.leggauss(24)


def _hankel_trunc_panels(radial, k, rmax):
    """int_0^rmax radial(r) J0(k r) r dr by composite 24-point Gauss-Legendre (numpy): panels a
    quarter period of J0 wide (at most rmax / 16), the first one refined geometrically towards
    r = 0 for the r^(1/n) cusp of the Sersic profile.  Converges to ~1e-15 of the k = 0 value."""
    w = min(0.5 * math.pi / k, rmax / 16.) if k > 0 else rmax / 16.
    edges = np.arange(0., rmax, w)
    edges = np.append(edges, rmax)
    fine = edges[1] * np.logspace(-40, 0, 41, base=2.)
    edges = np.concatenate([[0.], fine, edges[2:]])
    a, b = edges[:-1], edges[1:]
    x, wt = _GL24
    r = 0.5 * (a + b)[:, None] + 0.5 * (b - a)[:, None] * x[None, :]
    f = r * radial(r) * sps.j0(k * r)
    return float(np.sum(0.5 * (b - a) * (f @ wt)))


_SERSIC_INFO = {}


class SersicInfo:
    """SersicInfo (src/SBSersic.cpp:223-537), in units of r0; trunc > 0 (units of r0) for a
    truncated profile: the flux fraction P(2n, trunc^(1/n)), the truncated gamma functions of
    the low-k series, the stepK limit at trunc, and the table's Hankel transforms over
    [0, trunc] (here by QUADPACK to 1e-14; the reference's adaptive GKP stops at its own
    tolerance, ~3e-9 of the k = 0 value).

    Pinned against oracle/_ref (the reference's own SersicInfo compiled from its sources):
    table, ksq_min/ksq_max, high-k coefficients, maxK, stepK and HLR, truncated or not; and by
    the Sersic maxk/stepk/r0 known answers of tests/test_sersic.py:497-558 (n=3.1, truncated
    and untruncated).
    """

    def __init__(self, n, gsp, trunc=0.):
        self.n = float(n)
        self.gsp = gsp
        self.trunc = float(trunc)
        self.invn = 1. / self.n
        self.inv2n = 0.5 * self.invn
        self.gamma2n = math.gamma(2. * self.n)
        # getFluxFraction (src/SBSersic.cpp:335-350): SersicIntegratedFlux(n, trunc)
        self.flux_fraction = (_gamma_p(2. * self.n, fast_pow(self.trunc, self.invn))
                              if self.trunc > 0. else 1.)
        self._b = None
        self._build_ft()
        self._stepk = None

    @staticmethod
    def get(n, gsp, trunc=0.):
        key = (float(n), float(trunc), gsp.key())
        if key not in _SERSIC_INFO:
            _SERSIC_INFO[key] = SersicInfo(n, gsp, trunc)
        return _SERSIC_INFO[key]

    @property
    def b(self):
        if self._b is None:
            self._b = sersic_calculate_b(self.n, self.flux_fraction)
        return self._b

    @property
    def hlr(self):
        return math.pow(self.b, self.n)

    def missing_flux_radius(self, frac):
        """calculateMissingFluxRadius (src/SBSersic.cpp:558-625)."""
        n = self.n
        missing_flux = frac * self.gamma2n
        z1 = -math.log(missing_flux)
        if n == 0.5:
            z = z1
        else:
            z = 4. * (n + 1.)
            twonm1 = 2. * n - 1.
            z2 = z1 + twonm1 * math.log(z) + twonm1 / z + twonm1 * (2. * n - 3.) / (2. * z * z)
            if z2 > z1 and z2 - z1 < 0.01:
                z2 = z1 + 0.01
            elif z2 < z1 and z2 - z1 > -0.01:
                z2 = z1 - 0.01
            if z1 < 0.:
                z1 = self.b
            s = nm.Solve(_SersicMissingFlux(n, missing_flux), z1, z2)
            s.bracket_upper()
            z = s.root()
        return math.pow(z, n)

    @property
    def stepk(self):
        # SersicInfo::stepK (src/SBSersic.cpp:304-318)
        if self._stepk is None:
            R = self.missing_flux_radius(self.gsp.folding_threshold)
            if self.trunc > 0. and self.trunc < R:
                R = self.trunc
            R = max(R, self.gsp.stepk_minimum_hlr)
            self._stepk = math.pi / R
        return self._stepk

    def _build_ft(self):
        # SersicInfo::buildFT (src/SBSersic.cpp:387-537)
        gsp = self.gsp
        n = self.n
        gamma4n = math.gamma(4. * n)
        gamma6n = math.gamma(6. * n)
        gamma8n = math.gamma(8. * n)
        if self.trunc > 0.:
            z = math.pow(self.trunc, 1. / n)
            gamma4n *= _gamma_p(4. * n, z)
            gamma6n *= _gamma_p(6. * n, z)
            gamma8n *= _gamma_p(8. * n, z)
        ff = self.flux_fraction
        self.kderiv2 = -gamma4n / (4. * self.gamma2n) / ff
        self.kderiv4 = gamma6n / (64. * self.gamma2n) / ff
        kderiv6 = gamma8n / (2304. * self.gamma2n) / ff
        kmin = math.pow(gsp.kvalue_accuracy / kderiv6, 1. / 6.)
        self.ksq_min = kmin * kmin
        hankel_norm = ff * n * self.gamma2n
        dlogk = gsp.table_spacing * math.sqrt(math.sqrt(gsp.kvalue_accuracy / 10.))
        n_correct = 0
        n_fit = 10
        fit_vals = deque()
        sf = skf = sk = sk2 = 0.
        self.ksq_max = -1.
        maxk = kmin
        invn = self.invn
        radial = lambda r: nm.expd(-fast_pow(r, invn))
        found_maxk = False
        args, vals = [], []
        highk_a = highk_b = 0.
        logk = math.log(kmin) - 0.001
        while logk < math.log(500.):
            k = nm.expd(logk)
            ksq = k * k
            if self.trunc > 0.:
                val = _hankel_trunc_panels(radial, k, self.trunc)
            else:
                val = nm.hankel_inf(radial, k, gsp.integration_relerr,
                                    gsp.integration_abserr * hankel_norm)
            val /= hankel_norm
            f0 = val * ksq
            args.append(logk)
            vals.append(f0)
            if abs(val) > gsp.maxk_threshold:
                maxk = k
                n_correct = 0
            else:
                found_maxk = True
                highk_a = (sf * sk2 - sk * skf) / (n_fit * sk2 - sk * sk)
                highk_b = (n_fit * skf - sk * sf) / (n_fit * sk2 - sk * sk)
                f0_pred = highk_a + highk_b / k
                if abs(f0 - f0_pred) / ksq < gsp.kvalue_accuracy:
                    n_correct += 1
                else:
                    n_correct = 0
                if n_correct >= 5:
                    self.ksq_max = ksq
                    break
            if len(fit_vals) == n_fit:
                k_back = nm.expd(logk - n_fit * dlogk)
                f_back = fit_vals.pop()
                inv_k = 1. / k
                inv_k_back = 1. / k_back
                sf += f0 - f_back
                skf += f0 * inv_k - f_back * inv_k_back
                sk += inv_k - inv_k_back
                sk2 += inv_k * inv_k - inv_k_back * inv_k_back
            else:
                inv_k = 1. / k
                sf += f0
                skf += f0 * inv_k
                sk += inv_k
                sk2 += inv_k * inv_k
            fit_vals.appendleft(f0)
            logk += dlogk
        self.table = nm.SplineTable(args, vals)
        if self.ksq_max <= 0.:
            self.ksq_max = nm.expd(2. * self.table.arg_max)
        if found_maxk:
            maxk *= nm.expd(dlogk)
        else:
            highk_a = (sf * sk2 - sk * skf) / (n_fit * sk2 - sk * sk)
            highk_b = (n_fit * skf - sk * sf) / (n_fit * sk2 - sk * sk)
            maxk = math.sqrt(highk_a / gsp.maxk_threshold)
            for _ in range(3):
                maxk = math.sqrt((highk_a - highk_b / maxk) / gsp.maxk_threshold)
        self.highk_a = highk_a
        self.highk_b = highk_b
        self.maxk = maxk

    def kvalue(self, ksq):
        """SersicInfo::kValue (src/SBSersic.cpp:363-376)."""
        ksq = np.asarray(ksq, dtype=float)
        with np.errstate(divide='ignore', invalid='ignore'):
            hi = (self.highk_a + self.highk_b / np.sqrt(ksq)) / ksq
            lo = 1. + ksq * (self.kderiv2 + ksq * self.kderiv4)
            mid = self.table(0.5 * np.log(ksq)) / ksq
        return np.where(ksq >= self.ksq_max, hi, np.where(ksq < self.ksq_min, lo, mid))


class Sersic:
    """SBSersic (src/SBSersic.cpp:71-90, 237-281); galsim/sersic.py:201-250 (the truncation
    radius from the half-light radius, flux_untruncated)."""

    def __init__(self, n, half_light_radius=None, scale_radius=None, flux=1., trunc=0.,
                 flux_untruncated=False, gsparams=None):
        self.n = float(n)
        self.flux = float(flux)
        self.trunc = float(trunc)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        if half_light_radius is not None:
            if self.trunc == 0. or flux_untruncated:
                scale_radius = half_light_radius / sersic_hlr(self.n, 1.)
            else:
                scale_radius = sersic_truncated_scale(self.n, half_light_radius, self.trunc)
        self.r0 = float(scale_radius)
        self.info = SersicInfo.get(self.n, self.gsparams, self.trunc / self.r0)
        if self.trunc > 0. and flux_untruncated:
            self.flux *= self.info.flux_fraction

    @property
    def half_light_radius(self):
        return self.r0 * self.info.hlr

    @property
    def maxk(self):
        return self.info.maxk / self.r0

    @property
    def stepk(self):
        return self.info.stepk / self.r0

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        r0 = self.r0
        kx = _axis(kx0 * r0, dkx * r0, m)
        ky = _axis(ky0 * r0, dky * r0, n)
        return self.flux * self.info.kvalue(kx[None, :] * kx[None, :] + (ky * ky)[:, None])

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        r0 = self.r0
        kx, ky = _grid2d(kx0 * r0, dkx * r0, dkxy * r0, ky0 * r0, dky * r0, dkyx * r0, m, n)
        return self.flux * self.info.kvalue(kx * kx + ky * ky)


# ----------------------------------------------------------------------------------------------
# Box / Pixel -- src/SBBox.cpp, galsim/box.py
# ----------------------------------------------------------------------------------------------
class Box:
    """SBBox (src/SBBox.cpp:151-224); maxk/stepk galsim/box.py:100-105."""

    def __init__(self, width, height, flux=1., gsparams=None):
        self.width = float(width)
        self.height = float(height)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        self.wo2pi = self.width / (2. * math.pi)
        self.ho2pi = self.height / (2. * math.pi)

    @property
    def maxk(self):
        return 2. / (self.gsparams.maxk_threshold * min(self.width, self.height))

    @property
    def stepk(self):
        return math.pi / max(self.width, self.height)

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        sx = sinc(_axis(kx0 * self.wo2pi, dkx * self.wo2pi, m))
        sy = sinc(_axis(ky0 * self.ho2pi, dky * self.ho2pi, n))
        return (self.flux * sx[None, :]) * sy[:, None]

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        kx = _grid2d(kx0 * self.wo2pi, dkx * self.wo2pi, dkxy * self.wo2pi,
                     0., 0., 0., m, n)[0]
        ky = _grid2d(0., 0., 0., ky0 * self.ho2pi, dky * self.ho2pi, dkyx * self.ho2pi, m, n)[1]
        return (self.flux * sinc(kx)) * sinc(ky)


def Pixel(scale, flux=1., gsparams=None):
    """galsim/box.py:136-160."""
    return Box(scale, scale, flux, gsparams)


# ----------------------------------------------------------------------------------------------
# InclinedSersic -- src/SBInclinedSersic.cpp, galsim/inclined.py
# ----------------------------------------------------------------------------------------------
class InclinedSersic:
    """SBInclinedSersic (src/SBInclinedSersic.cpp:86-209 ctor, 227-309 fills, 321-370
    kValueHelper); galsim/inclined.py:226-475.  The Sersic k-value (SersicInfo) at
    (kx, ky cos i) times the vertical factor of InclinedExponential; maxk by Brent and the
    clipping radius by bisection on the k-value along the axes.

    Pinned by Lance Miller's independent images (n = 1 and the Sersic set of
    tests/test_inclined.py:58-63) and by n = 1 matching the InclinedExponential.
    """

    def __init__(self, n, inclination, scale_radius, scale_height, flux=1., trunc=0.,
                 gsparams=None):
        self.n = float(n)
        self.r0 = float(scale_radius)
        self.h0 = float(scale_height)
        self.flux = float(flux)
        self.trunc = float(trunc)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        gsp = self.gsparams
        self.info = SersicInfo.get(self.n, gsp, self.trunc / self.r0)
        self.cosi = abs(math.cos(inclination))
        self.hs = 0.5 * math.pi * self.h0 * abs(math.sin(inclination)) / self.r0
        kderiv6 = 31. / 15120. * self.hs
        self.ksq_min = (gsp.kvalue_accuracy / kderiv6) ** (1. / 3.) if kderiv6 > 0 else math.inf
        self.ksq_max = math.inf
        maxk_min = gsp.maxk_threshold ** (-1. / 3.)
        clipk_min = gsp.kvalue_accuracy ** (-1. / 3.)
        if self.cosi > 0.01:
            if self.cosi < 0.96:
                maxk_max, clipk_max = maxk_min / self.cosi, clipk_min / self.cosi
            else:
                maxk_max, clipk_max = 1.05 * maxk_min, 1.05 * clipk_min
        else:
            maxk_max, clipk_max = 100 * maxk_min, 100 * clipk_min

        def func(target):
            def f(k):
                return max(float(self.helper(k, 0.)), float(self.helper(0., k))) - target
            return f

        f = func(gsp.maxk_threshold)
        s = nm.Solve(f, maxk_min, maxk_max)
        if f(maxk_min) <= 0:
            s.bracket_lower_with_limit(0.)
        else:
            s.bracket_upper()
        self._maxk = s.root() + s.xtol
        f = func(gsp.kvalue_accuracy)
        s = nm.Solve(f, clipk_min, clipk_max)
        if f(clipk_min) <= 0:
            s.bracket_lower_with_limit(0.)
        else:
            s.bracket_upper()
        clipk = s.bisect() + s.xtol
        self.ksq_max = clipk * clipk

    def helper(self, kx, ky):
        kx = np.asarray(kx, dtype=float)
        ky = np.asarray(ky, dtype=float)
        ky_cosi = ky * self.cosi
        ksq = kx * kx + ky_cosi * ky_cosi
        base = self.info.kvalue(np.where(ksq > self.ksq_max, 1., ksq))
        sky = self.hs * ky
        sky2 = sky * sky
        with np.errstate(invalid='ignore', divide='ignore', over='ignore'):
            conv = np.where(sky2 < self.ksq_min,
                            1. - 0.16666666667 * sky2 * (1. - 0.116666666667 * sky2),
                            sky / np.sinh(sky))
        return np.where(ksq > self.ksq_max, 0., base * conv)

    @property
    def maxk(self):
        return self._maxk / self.r0

    @property
    def stepk(self):
        return self.info.stepk / self.r0

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; ky0 *= r0; dky *= r0
        kx = _axis(kx0, dkx, m)
        ky = _axis(ky0, dky, n)
        return self.flux * self.helper(kx[None, :], ky[:, None] + 0. * kx[None, :])

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; dkxy *= r0; ky0 *= r0; dky *= r0; dkyx *= r0
        out = np.zeros((n, m))
        rx = _axis(kx0, dkxy, n)
        ry = _axis(ky0, dky, n)
        for j in range(n):
            out[j] = self.flux * self.helper(_axis(rx[j], dkx, m), _axis(ry[j], dkyx, m))
        return out


# ----------------------------------------------------------------------------------------------
# TopHat and DeltaFunction -- src/SBBox.cpp:283-445, galsim/box.py, galsim/deltafunction.py
# ----------------------------------------------------------------------------------------------
class TopHat:
    """SBTopHat (src/SBBox.cpp:283-445): flux 2 J1(k r)/(k r) with its Taylor branch.  Pinned by
    the reference's known maxk / stepk (tests/test_box.py:275-297) and by the closed-form 2-d
    transform of a uniform disk integrated numerically (tests/test_tophat_cpu.py)."""

    def __init__(self, radius, flux=1., gsparams=None):
        self.r0 = float(radius)
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS

    @property
    def maxk(self):
        return (2. * math.sqrt(2. / math.pi) / self.gsparams.maxk_threshold) ** (2. / 3.) / self.r0

    @property
    def stepk(self):
        return math.pi / self.r0

    def kvalue2(self, kr0sq):
        kr0sq = np.asarray(kr0sq, dtype=float)
        kr0 = np.sqrt(kr0sq)
        with np.errstate(divide='ignore', invalid='ignore'):
            big = 2. * self.flux * sps.j1(kr0) / kr0
        small = self.flux * (1. - kr0sq * ((1. / 8.) + (1. / 192.) * kr0sq))
        return np.where(kr0sq < 1.e-4, small, big)

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        r0 = self.r0
        kx = _axis(kx0 * r0, dkx * r0, m)
        ky = _axis(ky0 * r0, dky * r0, n)
        return self.kvalue2(kx[None, :] * kx[None, :] + (ky * ky)[:, None])

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        r0 = self.r0
        kx0 *= r0; dkx *= r0; dkxy *= r0; ky0 *= r0; dky *= r0; dkyx *= r0
        out = np.zeros((n, m))
        rx = _axis(kx0, dkxy, n)
        ry = _axis(ky0, dky, n)
        for j in range(n):
            kx = _axis(rx[j], dkx, m)
            ky = _axis(ry[j], dkyx, m)
            out[j] = self.kvalue2(kx * kx + ky * ky)
        return out


class DeltaFunction:
    """galsim/deltafunction.py: the constant flux; maxk = stepk = 1e300."""

    def __init__(self, flux=1., gsparams=None):
        self.flux = float(flux)
        self.gsparams = gsparams or DEFAULT_GSPARAMS
        self.maxk = self.stepk = 1.e300

    def fill_1d(self, kx0, dkx, m, ky0, dky, n):
        return np.full((n, m), self.flux)

    def fill_2d(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        return np.full((n, m), self.flux)

