"""Kolmogorov (galsim/kolmogorov.py; src/SBKolmogorov.cpp)."""
import math

import numpy as np

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams

_KOLM_CACHE = {}


class Kolmogorov(Leaf):
    _fwhm_factor = 0.9758634299
    _hlr_factor = 0.5548101137
    _k0_factor = 2.992939911888651
    _leaf_type = _lib.LEAF_KOLMOGOROV
    _is_analytic_x = True

    def __init__(self, lam_over_r0=None, fwhm=None, half_light_radius=None, flux=1.,
                 gsparams=None):
        given = [v is not None for v in (lam_over_r0, fwhm, half_light_radius)]
        if sum(given) != 1:
            raise ValueError("One of lam_over_r0, fwhm, or half_light_radius must be specified")
        if fwhm is not None:
            lam_over_r0 = float(fwhm) / Kolmogorov._fwhm_factor
        elif half_light_radius is not None:
            lam_over_r0 = float(half_light_radius) / Kolmogorov._hlr_factor
        self._lor0 = float(lam_over_r0)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)
        self._k0 = Kolmogorov._k0_factor / self._lor0

    @property
    def lam_over_r0(self):
        return self._lor0

    @property
    def fwhm(self):
        return self._lor0 * Kolmogorov._fwhm_factor

    @property
    def half_light_radius(self):
        return self._lor0 * Kolmogorov._hlr_factor

    @property
    def _maxk(self):
        # galsim/kolmogorov.py:281-284
        return (-math.log(self.gsparams.kvalue_accuracy)) ** 0.6 * self._k0

    @property
    def _stepk(self):
        key = self.gsparams
        if key not in _KOLM_CACHE:
            out = np.zeros(2)
            _lib.check(_lib.lib().gsmi_kolmogorov_info(_lib.ptr(_lib.gsp_array(key)),
                                                       _lib.ptr(out)), "Kolmogorov info")
            _KOLM_CACHE[key] = out
        return float(_KOLM_CACHE[key][1]) * self._k0

    def _fill(self, rec, ctx, jac):
        rec['scale'] = 1. / self._k0
        rec['flux'] = self._flux

    def withFlux(self, flux):
        return Kolmogorov(lam_over_r0=self._lor0, flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Kolmogorov(lam_over_r0=%r, flux=%r)' % (self._lor0, self._flux)
