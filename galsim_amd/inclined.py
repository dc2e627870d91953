"""InclinedExponential (galsim/inclined.py; src/SBInclinedExponential.cpp): an exponential disk
with a sech^2 vertical profile seen at an inclination, in k-space."""
import math

import numpy as np

from . import _lib
from ._leaf import Leaf
from .angle import Angle
from .exponential import Exponential
from .gsparams import GSParams


def inclined_exp_info(cosi, hs, gsparams):
    """(maxk, stepk, ksq_min, ksq_max) in units of the scale radius (the
    SBInclinedExponentialImpl constructor), from libgsmi's host numerics."""
    out = np.zeros(4)
    _lib.check(_lib.lib().gsmi_inclined_exp_info(float(cosi), float(hs),
                                                 _lib.ptr(_lib.gsp_array(gsparams)),
                                                 _lib.ptr(out)), "InclinedExponential info")
    return out


class InclinedExponential(Leaf):
    """galsim/inclined.py:31-223.  The disk lies along x; its projected minor axis is y."""
    _leaf_type = _lib.LEAF_INCLINED_EXP

    def __init__(self, inclination, half_light_radius=None, scale_radius=None,
                 scale_height=None, scale_h_over_r=None, flux=1., gsparams=None):
        if scale_radius is not None:
            if not scale_radius > 0.:
                raise _lib.GalSimAMDRangeError("scale_radius must be > 0.")
            if half_light_radius is not None:
                raise ValueError("Only one of scale_radius and half_light_radius may be "
                                 "specified")
            self._r0 = float(scale_radius)
        elif half_light_radius is not None:
            if not half_light_radius > 0.:
                raise _lib.GalSimAMDRangeError("half_light_radius must be > 0.")
            self._r0 = float(half_light_radius) / Exponential._hlr_factor
        else:
            raise ValueError("Either scale_radius or half_light_radius must be specified")
        if scale_height is not None:
            if not scale_height > 0.:
                raise _lib.GalSimAMDRangeError("scale_height must be > 0.")
            if scale_h_over_r is not None:
                raise ValueError("Only one of scale_height and scale_h_over_r may be specified.")
            self._h0 = float(scale_height)
        else:
            if scale_h_over_r is None:
                scale_h_over_r = 0.1
            elif not scale_h_over_r > 0.:
                raise _lib.GalSimAMDRangeError("scale_h_over_r must be > 0.")
            self._h0 = float(self._r0) * float(scale_h_over_r)
        if not isinstance(inclination, Angle):
            raise TypeError("Input inclination should be an Angle")
        self._inclination = inclination
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)
        self._info_cache = None

    def _info(self):
        if self._info_cache is None:
            rad = self._inclination.rad
            cosi = abs(math.cos(rad))
            hs = 0.5 * math.pi * self._h0 * abs(math.sin(rad)) / self._r0
            self._info_cache = (cosi, hs, inclined_exp_info(cosi, hs, self.gsparams))
        return self._info_cache

    def _reset(self):
        self._info_cache = None

    @property
    def inclination(self):
        return self._inclination

    @property
    def scale_radius(self):
        return self._r0

    @property
    def scale_height(self):
        return self._h0

    @property
    def scale_h_over_r(self):
        return self._h0 / self._r0

    @property
    def disk_half_light_radius(self):
        return self._r0 * Exponential._hlr_factor

    @property
    def _maxk(self):
        return self._info()[2][0] / self._r0

    @property
    def _stepk(self):
        return self._info()[2][1] / self._r0

    def _fill(self, rec, ctx, jac):
        cosi, hs, info = self._info()
        rec['scale'] = self._r0
        rec['flux'] = self._flux
        rec['p'][0] = cosi
        rec['p'][1] = hs
        rec['p'][2] = info[2]
        rec['p'][3] = info[3]

    def withFlux(self, flux):
        return InclinedExponential(self._inclination, scale_radius=self._r0,
                                   scale_height=self._h0, flux=flux, gsparams=self.gsparams)

    def __eq__(self, o):
        return isinstance(o, InclinedExponential) and \
            (self._inclination, self._r0, self._h0, self._flux, self.gsparams) == \
            (o._inclination, o._r0, o._h0, o._flux, o.gsparams)

    def __hash__(self):
        return hash(('InclinedExponential', self._inclination, self._r0, self._h0, self._flux,
                     self.gsparams))

    def __repr__(self):
        return ('galsim_amd.InclinedExponential(inclination=%r, scale_radius=%r, '
                'scale_height=%r, flux=%r)' % (self._inclination, self._r0, self._h0,
                                              self._flux))
