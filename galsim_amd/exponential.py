"""Exponential (galsim/exponential.py; src/SBExponential.cpp)."""
import math

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams


class Exponential(Leaf):
    _hlr_factor = 1.6783469900166605
    _leaf_type = _lib.LEAF_EXPONENTIAL
    _is_analytic_x = True

    def __init__(self, half_light_radius=None, scale_radius=None, flux=1., gsparams=None):
        if half_light_radius is not None:
            if scale_radius is not None:
                raise ValueError("Only one of scale_radius and half_light_radius may be specified")
            scale_radius = half_light_radius / Exponential._hlr_factor
        elif scale_radius is None:
            raise ValueError("Either scale_radius or half_light_radius must be specified")
        self._r0 = float(scale_radius)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)

    @property
    def scale_radius(self):
        return self._r0

    @property
    def half_light_radius(self):
        return self._r0 * Exponential._hlr_factor

    @property
    def _maxk(self):
        # galsim/exponential.py:131-132
        return (self.gsparams.maxk_threshold ** (-1. / 3.)) / self._r0

    @property
    def _stepk(self):
        # ExponentialInfo::ExponentialInfo (src/SBExponential.cpp:527-544)
        logx = math.log(self.gsparams.folding_threshold)
        R = -logx
        for _ in range(3):
            R = math.log(1. + R) - logx
        R = max(R, self.gsparams.stepk_minimum_hlr * 1.6783469900166605)
        return (math.pi / R) / self._r0

    def _fill(self, rec, ctx, jac):
        kva = self.gsparams.kvalue_accuracy
        ksq_max = kva ** (-1. / 1.5) - 1.
        rec['scale'] = self._r0
        rec['flux'] = self._flux
        rec['p'][0] = ksq_max
        rec['p'][1] = math.sqrt(ksq_max)
        rec['cliprows'] = ctx.exponential_cliprows(self._r0, self.gsparams, jac)

    def withFlux(self, flux):
        return Exponential(scale_radius=self._r0, flux=flux, gsparams=self.gsparams)

    def __eq__(self, o):
        return isinstance(o, Exponential) and (self._r0, self._flux, self.gsparams) == \
            (o._r0, o._flux, o.gsparams)

    def __hash__(self):
        return hash(('Exponential', self._r0, self._flux, self.gsparams))

    def __repr__(self):
        return 'galsim_amd.Exponential(scale_radius=%r, flux=%r)' % (self._r0, self._flux)
