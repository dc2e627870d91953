"""Gaussian (galsim/gaussian.py; src/SBGaussian.cpp)."""
import math

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams


class Gaussian(Leaf):
    _fwhm_factor = 2.3548200450309493
    _hlr_factor = 1.1774100225154747
    _leaf_type = _lib.LEAF_GAUSSIAN
    _is_analytic_x = True

    def __init__(self, half_light_radius=None, sigma=None, fwhm=None, flux=1., gsparams=None):
        given = [v is not None for v in (half_light_radius, sigma, fwhm)]
        if sum(given) != 1:
            raise ValueError("One of sigma, fwhm, and half_light_radius must be specified")
        if fwhm is not None:
            sigma = fwhm / Gaussian._fwhm_factor
        elif half_light_radius is not None:
            sigma = half_light_radius / Gaussian._hlr_factor
        self._sigma = float(sigma)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)

    @property
    def sigma(self):
        return self._sigma

    @property
    def _maxk(self):
        return math.sqrt(-2. * math.log(self.gsparams.maxk_threshold)) / self._sigma

    @property
    def _stepk(self):
        R = max(math.sqrt(-2. * math.log(self.gsparams.folding_threshold)),
                self.gsparams.stepk_minimum_hlr * Gaussian._hlr_factor)
        return math.pi / (R * self._sigma)

    def _fill(self, rec, ctx, jac):
        kva = self.gsparams.kvalue_accuracy
        rec['scale'] = self._sigma
        rec['flux'] = self._flux
        rec['p'][0] = -2. * math.log(kva)                 # SBGaussian.cpp:62
        rec['p'][1] = (kva * 48.) ** (1. / 3.)            # SBGaussian.cpp:67

    def withFlux(self, flux):
        return Gaussian(sigma=self._sigma, flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Gaussian(sigma=%r, flux=%r)' % (self._sigma, self._flux)
