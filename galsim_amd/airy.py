"""Airy (galsim/airy.py; src/SBAiry.cpp)."""
import math

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams

_ANGLE_UNITS = {'radians': 1., 'rad': 1., 'degrees': math.pi / 180., 'deg': math.pi / 180.,
                'hours': math.pi / 12., 'hour': math.pi / 12.,
                'arcmin': math.pi / (180. * 60.), 'arcsec': math.pi / (180. * 3600.)}


class Airy(Leaf):
    """The diffraction pattern of a circular (optionally obscured) aperture: in k-space the
    overlap area of two annuli (galsim/airy.py:28-249, src/SBAiry.cpp)."""
    _leaf_type = _lib.LEAF_AIRY
    _is_analytic_x = True
    _hlr_factor = 0.5348321477242647
    _fwhm_factor = 1.028993969962188

    def __init__(self, lam_over_diam=None, lam=None, diam=None, obscuration=0., flux=1.,
                 scale_unit=None, gsparams=None):
        self._obscuration = float(obscuration)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)
        if lam_over_diam is not None:
            if lam is not None or diam is not None:
                raise ValueError("If specifying lam_over_diam, then do not specify lam or diam")
            self._lod = float(lam_over_diam)
        else:
            if lam is None or diam is None:
                raise ValueError("If not specifying lam_over_diam, then specify lam AND diam")
            unit = 'arcsec' if scale_unit is None else scale_unit
            if isinstance(unit, str):
                unit = _ANGLE_UNITS[unit]
            else:
                unit = float(getattr(unit, 'value', unit))
            self._lod = (1.e-9 * float(lam) / float(diam)) / unit

    @property
    def lam_over_diam(self):
        return self._lod

    @property
    def obscuration(self):
        return self._obscuration

    @property
    def half_light_radius(self):
        if self._obscuration == 0.:
            return self._lod * Airy._hlr_factor
        raise NotImplementedError("Half light radius calculation not implemented for Airy "
                                  "objects with non-zero obscuration.")

    @property
    def fwhm(self):
        if self._obscuration == 0.:
            return self._lod * Airy._fwhm_factor
        raise NotImplementedError("FWHM calculation not implemented for Airy objects with "
                                  "non-zero obscuration.")

    @property
    def _maxk(self):
        return 2. * math.pi / self._lod

    @property
    def _stepk(self):
        # AiryInfoObs / AiryInfoNoObs ctors (src/SBAiry.cpp:332-348, 445-455), times D
        R = 1. / (self.gsparams.folding_threshold * 0.5 * math.pi * math.pi *
                  (1. - self._obscuration))
        R = max(R, self.gsparams.stepk_minimum_hlr * 0.5348321477)
        return (math.pi / R) / self._lod

    def _fill(self, rec, ctx, jac):
        obs = self._obscuration
        rec['scale'] = self._lod / math.pi
        rec['flux'] = self._flux / (math.pi * (1. - obs * obs))
        rec['p'][0] = obs
        rec['p'][1] = obs * obs

    def withFlux(self, flux):
        return Airy(lam_over_diam=self._lod, obscuration=self._obscuration, flux=flux,
                    gsparams=self.gsparams)

    def __eq__(self, o):
        return isinstance(o, Airy) and (self._lod, self._obscuration, self._flux,
                                        self.gsparams) == (o._lod, o._obscuration, o._flux,
                                                           o.gsparams)

    def __hash__(self):
        return hash(('Airy', self._lod, self._obscuration, self._flux, self.gsparams))

    def __repr__(self):
        return 'galsim_amd.Airy(lam_over_diam=%r, obscuration=%r, flux=%r)' % (
            self._lod, self._obscuration, self._flux)
