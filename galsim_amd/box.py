"""Box and Pixel (galsim/box.py; src/SBBox.cpp)."""
import math

from . import _lib
from ._leaf import Leaf
from .gsparams import GSParams


class Box(Leaf):
    _leaf_type = _lib.LEAF_BOX
    _has_hard_edges = True
    _is_analytic_x = True

    def __init__(self, width, height, flux=1., gsparams=None):
        self._width = float(width)
        self._height = float(height)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)

    @property
    def width(self):
        return self._width

    @property
    def height(self):
        return self._height

    @property
    def _maxk(self):
        return 2. / (self.gsparams.maxk_threshold * min(self._width, self._height))

    @property
    def _stepk(self):
        return math.pi / max(self._width, self._height)

    def _fill(self, rec, ctx, jac):
        rec['scale'] = self._width / (2. * math.pi)
        rec['scale2'] = self._height / (2. * math.pi)
        rec['flux'] = self._flux

    def withFlux(self, flux):
        return Box(self._width, self._height, flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Box(width=%r, height=%r, flux=%r)' % (self._width, self._height,
                                                                 self._flux)


class Pixel(Box):
    def __init__(self, scale, flux=1., gsparams=None):
        super().__init__(scale, scale, flux=flux, gsparams=gsparams)

    @property
    def scale(self):
        return self._width

    def withFlux(self, flux):
        return Pixel(self._width, flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.Pixel(scale=%r, flux=%r)' % (self._width, self._flux)
