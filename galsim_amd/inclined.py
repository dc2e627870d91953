"""InclinedExponential (galsim/inclined.py; src/SBInclinedExponential.cpp): an exponential disk
with a sech^2 vertical profile seen at an inclination, in k-space."""
import math

import numpy as np

from . import _lib
from ._leaf import Leaf, jac_record
from .angle import Angle
from .exponential import Exponential
from .gsobject import GSObject
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


def inclined_sersic_info(n, trunc, cosi, hs, gsparams):
    """(maxk in 1/r0, ksq_min, ksq_max) of the SBInclinedSersicImpl constructor."""
    out = np.zeros(3)
    _lib.check(_lib.lib().gsmi_inclined_sersic_info(float(n), float(trunc), float(cosi),
                                                    float(hs),
                                                    _lib.ptr(_lib.gsp_array(gsparams)),
                                                    _lib.ptr(out)), "InclinedSersic info")
    return out


class InclinedSersic(GSObject):
    """galsim/inclined.py:226-475: a Sersic disk with a sech^2 vertical profile, inclined.  In
    k-space the Sersic value at (kx, ky cos i) times the vertical factor; lowered as a Sersic
    leaf (Jacobian scaling ky' by cos i) multiplied by an INCLINED_VERT leaf."""

    def __init__(self, n, inclination, half_light_radius=None, scale_radius=None,
                 scale_height=None, scale_h_over_r=None, flux=1., trunc=0.,
                 flux_untruncated=False, gsparams=None):
        from .sersic import Sersic
        if trunc < 0.:
            raise _lib.GalSimAMDRangeError("trunc must be >= 0. (zero implying no truncation).")
        if scale_radius is not None and not scale_radius > 0.:
            raise _lib.GalSimAMDRangeError("scale_radius must be > 0.")
        if half_light_radius is not None and not half_light_radius > 0.:
            raise _lib.GalSimAMDRangeError("half_light_radius must be > 0.")
        # the radius and flux conventions are the Sersic's (galsim/inclined.py:330-367)
        self._base = Sersic(n, half_light_radius=half_light_radius, scale_radius=scale_radius,
                            flux=flux, trunc=trunc, flux_untruncated=flux_untruncated,
                            gsparams=gsparams)
        self._n = float(n)
        self._inclination = inclination
        self._trunc = float(trunc)
        self._gsparams = self._base.gsparams
        self._r0 = self._base.scale_radius
        self._flux = self._base.flux
        self._hlr = 0.
        if scale_height is not None:
            if not scale_height > 0.:
                raise _lib.GalSimAMDRangeError("scale_height must be > 0.")
            if scale_h_over_r is not None:
                raise ValueError("Only one of scale_height and scale_h_over_r may be specified")
            self._h0 = float(scale_height)
        else:
            if scale_h_over_r is None:
                scale_h_over_r = 0.1
            elif not scale_h_over_r > 0.:
                raise _lib.GalSimAMDRangeError("scale_h_over_r must be > 0.")
            self._h0 = float(scale_h_over_r) * self._r0
        if not isinstance(inclination, Angle):
            raise TypeError("Input inclination should be an Angle")
        self._info_cache = None

    def _info(self):
        if self._info_cache is None:
            rad = self._inclination.rad
            cosi = abs(math.cos(rad))
            hs = 0.5 * math.pi * self._h0 * abs(math.sin(rad)) / self._r0
            self._info_cache = (cosi, hs, inclined_sersic_info(self._n, self._trunc / self._r0,
                                                               cosi, hs, self.gsparams))
        return self._info_cache

    def withGSParams(self, gsparams=None, **kwargs):
        ret = GSObject.withGSParams(self, gsparams, **kwargs)
        if ret is not self:
            ret._base = self._base.withGSParams(ret._gsparams)
            ret._info_cache = None
        return ret

    @property
    def n(self):
        return self._n

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
    def trunc(self):
        return self._trunc

    @property
    def disk_half_light_radius(self):
        if self._hlr == 0.:
            self._hlr = self._base.half_light_radius
        return self._hlr

    @property
    def _maxk(self):
        return self._info()[2][0] / self._r0

    @property
    def _stepk(self):
        return self._base.stepk

    def _lower(self, ctx, jac):
        cosi, hs, info = self._info()
        j = np.eye(2) if jac is None else np.asarray(jac, dtype=float).reshape(2, 2)
        self._base._lower(ctx, j @ np.diag([1., cosi]))
        from .render import LeafRecord
        rec = LeafRecord()
        sep, jj = jac_record(jac)
        rec['type'] = _lib.LEAF_INCLINED_VERT
        rec['separable'] = sep
        rec['jac'] = jj
        rec['cliprows'] = -1
        rec['scale'] = self._r0
        rec['flux'] = 1.
        rec['p'][0] = cosi
        rec['p'][1] = hs
        rec['p'][2] = info[1]
        rec['p'][3] = info[2]
        ctx.emit(_lib.OP_LEAF, ctx.add_leaf(rec))
        ctx.emit(_lib.OP_MUL)

    def withFlux(self, flux):
        return InclinedSersic(self._n, self._inclination, scale_radius=self._r0,
                              scale_height=self._h0, flux=flux, trunc=self._trunc,
                              gsparams=self.gsparams)

    def __repr__(self):
        return ('galsim_amd.InclinedSersic(n=%r, inclination=%r, scale_radius=%r, '
                'scale_height=%r, flux=%r, trunc=%r)' % (self._n, self._inclination, self._r0,
                                                        self._h0, self._flux, self._trunc))
