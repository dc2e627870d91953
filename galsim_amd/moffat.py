"""Moffat, untruncated or truncated (galsim/moffat.py; src/SBMoffat.cpp)."""
import ctypes
import math

import numpy as np

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams


def temme_constants(nu):
    """The order-only constants of Temme's series for K_nu(x), x < 2 (N. M. Temme, J. Comput.
    Phys. 19, 324 (1975)), which the FP32 kernels take precomputed: with mu = nu - round(nu),
    fact = pi mu / sin(pi mu), gam1 = (1/Gamma(1-mu) - 1/Gamma(1+mu)) / (2 mu),
    gam2 = (1/Gamma(1-mu) + 1/Gamma(1+mu)) / 2, returns (fact gam1, fact gam2,
    Gamma(1+mu) / 2, Gamma(1-mu) / 2)."""
    nl = int(nu + 0.5)
    mu = nu - nl
    gampl = 1. / math.gamma(1. + mu)
    gammi = 1. / math.gamma(1. - mu)
    if abs(mu) < 1e-4:
        # Taylor coefficients of 1/Gamma(1+z) (Abramowitz & Stegun 6.1.34)
        m2 = mu * mu
        gam1 = -(0.5772156649015329 - 0.0420026350340952 * m2)
        gam2 = 1. - 0.6558780715202538 * m2 + 0.1665386113822915 * m2 * m2
        fact = 1. + (math.pi * mu) ** 2 / 6.
    else:
        gam1 = (gammi - gampl) / (2. * mu)
        gam2 = (gammi + gampl) / 2.
        fact = math.pi * mu / math.sin(math.pi * mu)
    return fact * gam1, fact * gam2, 0.5 / gampl, 0.5 / gammi


class Moffat(Leaf):
    _leaf_type = _lib.LEAF_MOFFAT
    _is_analytic_x = True

    @property
    def _has_hard_edges(self):
        # a truncated Moffat ends at trunc (galsim/moffat.py:203-204)
        return self._trunc != 0.

    def __init__(self, beta, scale_radius=None, half_light_radius=None, fwhm=None, trunc=0.,
                 flux=1., gsparams=None):
        self._beta = float(beta)
        self._trunc = float(trunc)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)
        if self._trunc == 0. and self._beta <= 1.1:
            raise _lib.GalSimAMDRangeError("Moffat profiles with beta <= 1.1 must be truncated")
        if self._trunc < 0.:
            raise _lib.GalSimAMDRangeError("Moffat trunc must be >= 0")
        given = [v is not None for v in (scale_radius, half_light_radius, fwhm)]
        if sum(given) != 1:
            raise ValueError("One of scale_radius, half_light_radius, or fwhm must be specified")
        if half_light_radius is not None:
            hlr = float(half_light_radius)
            out = np.zeros(1)
            if self._trunc > 0.:
                # galsim/moffat.py: trunc must exceed sqrt(2) hlr; MoffatCalculateSRFromHLR
                if self._trunc <= math.sqrt(2.) * hlr:
                    raise _lib.GalSimAMDRangeError(
                        "Moffat trunc must be > sqrt(2) * half_light_radius.")
                _lib.check(_lib.lib().gsmi_moffat_scale_radius_from_hlr_trunc(
                    hlr, self._trunc, self._beta, _lib.ptr(out)), "Moffat hlr")
            else:
                _lib.check(_lib.lib().gsmi_moffat_scale_radius_from_hlr(
                    hlr, self._beta, _lib.ptr(out)), "Moffat hlr")
            self._r0 = float(out[0])
        elif fwhm is not None:
            self._r0 = float(fwhm) / (2. * math.sqrt(2. ** (1. / self._beta) - 1.))
        else:
            self._r0 = float(scale_radius)
        self._info = None

    def _reset(self):
        self._info = None

    def _get_info(self):
        """Untruncated: (maxK, stepK, knorm, knorm2, kind).  Truncated: (maxK, stepK, flux
        factor, half-light radius, knots, knot spacing) and the kV_trunc table."""
        if self._info is None:
            g = _lib.ptr(_lib.gsp_array(self.gsparams))
            if self._trunc > 0.:
                out = np.zeros(6)
                _lib.check(_lib.lib().gsmi_moffat_trunc_info(self._beta, self._r0, self._trunc,
                                                             g, _lib.ptr(out)), "Moffat info")
                n = int(out[4])
                args, vals, y2 = np.zeros(n), np.zeros(n), np.zeros(n)
                _lib.check(_lib.lib().gsmi_moffat_trunc_table(
                    self._beta, self._r0, self._trunc, g, _lib.ptr(args), _lib.ptr(vals),
                    _lib.ptr(y2), n), "Moffat table")
                self._table = (args, vals, y2)
            else:
                out = np.zeros(5)
                _lib.check(_lib.lib().gsmi_moffat_info(self._beta, self._r0, self._flux, g,
                                                       _lib.ptr(out)), "Moffat info")
            self._info = out
        return self._info

    @property
    def trunc(self):
        return self._trunc

    @property
    def half_light_radius(self):
        if self._trunc > 0.:
            return float(self._get_info()[3])
        return self._r0 * math.sqrt(2. ** (1. / (self._beta - 1.)) - 1.)

    @property
    def beta(self):
        return self._beta

    @property
    def scale_radius(self):
        return self._r0

    @property
    def fwhm(self):
        return self._r0 * (2. * math.sqrt(2. ** (1. / self._beta) - 1.))

    @property
    def _maxk(self):
        return float(self._get_info()[0])

    @property
    def _stepk(self):
        return float(self._get_info()[1])

    def _fill(self, rec, ctx, jac):
        info = self._get_info()
        rec['scale'] = self._r0
        if self._trunc > 0.:
            # kV_trunc: knorm = flux (knorm2 = 1), the table on (k rD)^2
            args, vals, y2 = self._table
            rec['flux'] = self._flux
            rec['kind'] = 1
            rec['table'] = ctx.add_table(('moffat', self._beta, self._trunc / self._r0,
                                          self.gsparams), args, vals, y2)
            rec['ntab'] = len(args)
            rec['p'][0] = self._beta
            rec['p'][1] = 1. / float(info[5])
            return
        rec['flux'] = info[2]                 # knorm
        rec['kind'] = int(info[4])
        rec['p'][0] = self._beta
        rec['p'][1] = self._flux / info[2] if info[2] != 0 else 0.   # kV_gen at k = 0
        if int(info[4]) == 0:
            rec['p'][2:6] = temme_constants(self._beta - 1.)

    def withFlux(self, flux):
        return Moffat(self._beta, scale_radius=self._r0, trunc=self._trunc, flux=flux,
                      gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Moffat(beta=%r, scale_radius=%r, trunc=%r, flux=%r)' % (
            self._beta, self._r0, self._trunc, self._flux)
