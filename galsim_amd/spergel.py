"""Spergel (galsim/spergel.py; src/SBSpergel.cpp)."""
import ctypes

import numpy as np

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams


def spergel_info(nu, gsparams):
    """(maxk, stepk, hlr, ksq_max) in units of the scale radius (SpergelInfo,
    src/SBSpergel.cpp:529-653), from libgsmi's host numerics."""
    out = np.zeros(4)
    _lib.check(_lib.lib().gsmi_spergel_info(float(nu), _lib.ptr(_lib.gsp_array(gsparams)),
                                            _lib.ptr(out)), "Spergel info")
    return out


def _flux_radius(nu, f):
    out = ctypes.c_double()
    _lib.check(_lib.lib().gsmi_spergel_flux_radius(float(nu), float(f), ctypes.byref(out)),
               "Spergel flux radius")
    return out.value


class Spergel(Leaf):
    """flux (1 + (k r0)^2)^(-1-nu) in k-space (galsim/spergel.py:28-239)."""
    _leaf_type = _lib.LEAF_SPERGEL
    _is_analytic_x = True
    _minimum_nu = -0.85
    _maximum_nu = 4.0

    def __init__(self, nu, half_light_radius=None, scale_radius=None, flux=1., gsparams=None):
        self._nu = float(nu)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)
        if self._nu < Spergel._minimum_nu:
            raise _lib.GalSimAMDRangeError("Requested Spergel index is too small")
        if self._nu > Spergel._maximum_nu:
            raise _lib.GalSimAMDRangeError("Requested Spergel index is too large")
        if half_light_radius is not None:
            if scale_radius is not None:
                raise ValueError("Only one of scale_radius or half_light_radius may be specified")
            self._hlr = float(half_light_radius)
            self._r0 = self._hlr / _flux_radius(self._nu, 0.5)
        elif scale_radius is not None:
            self._r0 = float(scale_radius)
            self._hlr = 0.
        else:
            raise ValueError("Either scale_radius or half_light_radius must be specified for "
                             "Spergel")

    @property
    def nu(self):
        return self._nu

    @property
    def scale_radius(self):
        return self._r0

    @property
    def half_light_radius(self):
        if self._hlr == 0.:
            self._hlr = self._r0 * _flux_radius(self._nu, 0.5)
        return self._hlr

    def calculateIntegratedFlux(self, r):
        out = ctypes.c_double()
        _lib.check(_lib.lib().gsmi_spergel_integrated_flux(self._nu, float(r) / self._r0,
                                                           ctypes.byref(out)),
                   "Spergel integrated flux")
        return out.value

    def calculateFluxRadius(self, f):
        return _flux_radius(self._nu, f) * self._r0

    @property
    def _maxk(self):
        return spergel_info(self._nu, self.gsparams)[0] / self._r0

    @property
    def _stepk(self):
        return spergel_info(self._nu, self.gsparams)[1] / self._r0

    def _fill(self, rec, ctx, jac):
        ksq_max = spergel_info(self._nu, self.gsparams)[3]
        rec['scale'] = self._r0
        rec['flux'] = self._flux
        rec['p'][0] = ksq_max
        rec['p'][1] = -1. - self._nu
        rec['cliprows'] = ctx.kclip_rows(ksq_max, self._r0, jac)

    def withFlux(self, flux):
        return Spergel(self._nu, scale_radius=self._r0, flux=flux, gsparams=self.gsparams)

    def __eq__(self, o):
        return isinstance(o, Spergel) and (self._nu, self._r0, self._flux, self.gsparams) == \
            (o._nu, o._r0, o._flux, o.gsparams)

    def __hash__(self):
        return hash(('Spergel', self._nu, self._r0, self._flux, self.gsparams))

    def __repr__(self):
        return 'galsim_amd.Spergel(nu=%r, scale_radius=%r, flux=%r)' % (self._nu, self._r0,
                                                                         self._flux)
