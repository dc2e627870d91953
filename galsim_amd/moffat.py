"""Moffat, untruncated (galsim/moffat.py; src/SBMoffat.cpp)."""
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

    def __init__(self, beta, scale_radius=None, half_light_radius=None, fwhm=None, trunc=0.,
                 flux=1., gsparams=None):
        self._beta = float(beta)
        self._trunc = float(trunc)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)
        if self._trunc != 0.:
            raise NotImplementedError("truncated Moffat profiles are not part of galsim_amd yet")
        if self._beta <= 1.1:
            raise _lib.GalSimAMDRangeError("Moffat profiles with beta <= 1.1 must be truncated")
        given = [v is not None for v in (scale_radius, half_light_radius, fwhm)]
        if sum(given) != 1:
            raise ValueError("One of scale_radius, half_light_radius, or fwhm must be specified")
        if half_light_radius is not None:
            out = np.zeros(1)
            _lib.check(_lib.lib().gsmi_moffat_scale_radius_from_hlr(
                float(half_light_radius), self._beta, _lib.ptr(out)), "Moffat hlr")
            self._r0 = float(out[0])
        elif fwhm is not None:
            self._r0 = float(fwhm) / (2. * math.sqrt(2. ** (1. / self._beta) - 1.))
        else:
            self._r0 = float(scale_radius)
        self._info = None

    def _reset(self):
        self._info = None

    def _get_info(self):
        if self._info is None:
            out = np.zeros(5)
            _lib.check(_lib.lib().gsmi_moffat_info(self._beta, self._r0, self._flux,
                                                   _lib.ptr(_lib.gsp_array(self.gsparams)),
                                                   _lib.ptr(out)), "Moffat info")
            self._info = out
        return self._info

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
        rec['flux'] = info[2]                 # knorm
        rec['kind'] = int(info[4])
        rec['p'][0] = self._beta
        rec['p'][1] = self._flux / info[2] if info[2] != 0 else 0.   # kV_gen at k = 0
        if int(info[4]) == 0:
            rec['p'][2:6] = temme_constants(self._beta - 1.)

    def withFlux(self, flux):
        return Moffat(self._beta, scale_radius=self._r0, flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Moffat(beta=%r, scale_radius=%r, flux=%r)' % (
            self._beta, self._r0, self._flux)
