"""PositionD/PositionI and BoundsI (galsim/position.py, galsim/bounds.py), the subset the FFT
drawing path uses."""
import numpy as np


class PositionD:
    def __init__(self, x=0., y=0.):
        if isinstance(x, (PositionD, PositionI)):
            x, y = x.x, x.y
        elif isinstance(x, (tuple, list, np.ndarray)):
            x, y = x
        self.x = float(x)
        self.y = float(y)

    def __add__(self, o):
        return PositionD(self.x + o.x, self.y + o.y)

    def __sub__(self, o):
        return PositionD(self.x - o.x, self.y - o.y)

    def __eq__(self, o):
        return isinstance(o, (PositionD, PositionI)) and self.x == o.x and self.y == o.y

    def __repr__(self):
        return 'galsim_amd.PositionD(x=%r, y=%r)' % (self.x, self.y)


class PositionI(PositionD):
    def __init__(self, x=0, y=0):
        if isinstance(x, (PositionD, PositionI)):
            x, y = x.x, x.y
        self.x = int(x)
        self.y = int(y)

    def __repr__(self):
        return 'galsim_amd.PositionI(x=%r, y=%r)' % (self.x, self.y)


class BoundsI:
    """Integer bounds, inclusive (galsim/bounds.py:357-470)."""

    def __init__(self, xmin=None, xmax=None, ymin=None, ymax=None):
        if xmin is None:
            self._defined = False
            self.xmin = self.ymin = 0
            self.xmax = self.ymax = -1
            return
        self.xmin, self.xmax, self.ymin, self.ymax = int(xmin), int(xmax), int(ymin), int(ymax)
        self._defined = self.xmin <= self.xmax and self.ymin <= self.ymax

    def isDefined(self):
        return self._defined

    def includes(self, x, y):
        return self._defined and self.xmin <= x <= self.xmax and self.ymin <= y <= self.ymax

    @property
    def center(self):
        # galsim/bounds.py:446-455
        return PositionI(self.xmin + (self.xmax - self.xmin + 1) // 2,
                         self.ymin + (self.ymax - self.ymin + 1) // 2)

    @property
    def true_center(self):
        return PositionD((self.xmax + self.xmin) / 2., (self.ymax + self.ymin) / 2.)

    def numpyShape(self):
        if self._defined:
            return self.ymax - self.ymin + 1, self.xmax - self.xmin + 1
        return 0, 0

    def _getinitargs(self):
        return (self.xmin, self.xmax, self.ymin, self.ymax) if self._defined else ()

    def shift(self, delta):
        return BoundsI(self.xmin + delta.x, self.xmax + delta.x, self.ymin + delta.y,
                       self.ymax + delta.y)

    def __eq__(self, o):
        return isinstance(o, BoundsI) and self._getinitargs() == o._getinitargs()

    def __repr__(self):
        return 'galsim_amd.BoundsI(%r, %r, %r, %r)' % (self.xmin, self.xmax, self.ymin, self.ymax)
