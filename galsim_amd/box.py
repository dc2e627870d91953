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


class TopHat(Leaf):
    """A uniform disk of the given radius (galsim/box.py:177-260): flux 2 J1(k r) / (k r) in
    k-space (src/SBBox.cpp:283-445)."""
    _leaf_type = _lib.LEAF_TOPHAT
    _has_hard_edges = True
    _is_analytic_x = True

    def __init__(self, radius, flux=1., gsparams=None):
        self._radius = float(radius)
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)

    @property
    def radius(self):
        return self._radius

    @property
    def _maxk(self):
        # SBTopHatImpl::maxK (src/SBBox.cpp:430-436)
        return (2. * math.sqrt(2. / math.pi) / self.gsparams.maxk_threshold) ** (2. / 3.) / \
            self._radius

    @property
    def _stepk(self):
        return math.pi / self._radius

    def _fill(self, rec, ctx, jac):
        rec['scale'] = self._radius
        rec['flux'] = self._flux

    def withFlux(self, flux):
        return TopHat(self._radius, flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.TopHat(radius=%r, flux=%r)' % (self._radius, self._flux)


class DeltaFunction(Leaf):
    """A point source (galsim/deltafunction.py): the constant flux in k-space, with the
    reference's stand-in for infinite maxk and stepk."""
    _leaf_type = _lib.LEAF_DELTA
    _mock_inf = 1.e300

    def __init__(self, flux=1., gsparams=None):
        self._flux = float(flux)
        self._gsparams = GSParams.check(gsparams)

    @property
    def _maxk(self):
        return DeltaFunction._mock_inf

    @property
    def _stepk(self):
        return DeltaFunction._mock_inf

    def _fill(self, rec, ctx, jac):
        rec['flux'] = self._flux
        rec['scale'] = 1.

    def withFlux(self, flux):
        return DeltaFunction(flux=flux, gsparams=self.gsparams)

    def __repr__(self):
        return 'galsim_amd.DeltaFunction(flux=%r)' % (self._flux,)
