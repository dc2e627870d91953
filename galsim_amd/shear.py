"""Shear (galsim/shear.py:26-433): the g1/g2 and e1/e2 constructors and getMatrix."""
import numpy as np


class Shear:
    def __init__(self, *args, **kwargs):
        if args:
            if len(args) > 1 or kwargs:
                raise TypeError("Shear takes one complex argument or keywords")
            self._g = complex(args[0])
        elif 'g1' in kwargs or 'g2' in kwargs:
            self._g = kwargs.pop('g1', 0.) + 1j * kwargs.pop('g2', 0.)
            if abs(self._g) > 1.:
                raise ValueError("Requested shear exceeds 1.")
        elif 'e1' in kwargs or 'e2' in kwargs:
            e1 = kwargs.pop('e1', 0.)
            e2 = kwargs.pop('e2', 0.)
            absesq = e1 ** 2 + e2 ** 2
            if absesq > 1.:
                raise ValueError("Requested distortion exceeds 1.")
            self._g = (e1 + 1j * e2) * self._e2g(absesq)
        else:
            self._g = 0j
        if kwargs:
            raise TypeError("Shear got unexpected keyword arguments %s" % list(kwargs))

    @staticmethod
    def _e2g(absesq):
        # galsim/shear.py:324-330
        if absesq > 1.e-4:
            return 1. / (1. + np.sqrt(1. - absesq))
        return 0.5 + absesq * (0.125 + absesq * (0.0625 + absesq * 0.0390625))

    @property
    def g1(self):
        return self._g.real

    @property
    def g2(self):
        return self._g.imag

    @property
    def g(self):
        return abs(self._g)

    def getMatrix(self):
        # galsim/shear.py:365-386
        return np.array([[1. + self.g1, self.g2],
                         [self.g2, 1. - self.g1]]) / np.sqrt(1. - self.g ** 2)

    def __repr__(self):
        return 'galsim_amd.Shear(%r)' % self._g
