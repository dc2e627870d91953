"""Convolution (galsim/convolve.py), Fourier-space convolution only."""
import numpy as np

from . import _lib
from .gsobject import GSObject
from .gsparams import GSParams


def Convolve(*args, **kwargs):
    """galsim/convolve.py:32-69 for GSObject inputs."""
    if len(args) == 0:
        raise TypeError("At least one GSObject must be provided.")
    if len(args) == 1 and isinstance(args[0], (list, tuple)):
        args = args[0]
    return Convolution(*args, **kwargs)


class Convolution(GSObject):
    """galsim/convolve.py:71-426: the k-image is the product of the factors' k-images."""

    def __init__(self, *args, **kwargs):
        if len(args) == 1 and isinstance(args[0], (list, tuple)):
            args = args[0]
        if len(args) == 0:
            raise TypeError("At least one GSObject must be provided.")
        real_space = kwargs.pop("real_space", None)
        gsparams = kwargs.pop("gsparams", None)
        self._propagate_gsparams = kwargs.pop('propagate_gsparams', True)
        if kwargs:
            raise TypeError("Convolution got unexpected keyword argument(s): %s" % list(kwargs))
        for obj in args:
            if not isinstance(obj, GSObject):
                raise TypeError("Arguments to Convolution must be GSObjects, not %s" % obj)
        hard_edge = all(obj.has_hard_edges for obj in args)
        if real_space is None:
            real_space = hard_edge if len(args) <= 2 else False
        if real_space:
            raise NotImplementedError("real-space convolution is not part of galsim_amd")
        self._real_space = False
        if gsparams is None:
            self._gsparams = GSParams.combine([obj.gsparams for obj in args])
        else:
            self._gsparams = GSParams.check(gsparams)
        if self._propagate_gsparams:
            self._obj_list = [obj.withGSParams(self._gsparams) for obj in args]
        else:
            self._obj_list = list(args)

    @property
    def obj_list(self):
        return self._obj_list

    @property
    def real_space(self):
        return self._real_space

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        if self._propagate_gsparams:
            ret._obj_list = [obj.withGSParams(ret._gsparams) for obj in self.obj_list]
        return ret

    @property
    def _maxk(self):
        return min(obj.maxk for obj in self.obj_list)

    @property
    def _stepk(self):
        return sum(obj.stepk ** (-2) for obj in self.obj_list) ** (-0.5)

    @property
    def _flux(self):
        return float(np.prod([obj.flux for obj in self.obj_list]))

    @property
    def _has_hard_edges(self):
        return len(self.obj_list) == 1 and self.obj_list[0].has_hard_edges

    @property
    def _is_analytic_x(self):
        return len(self.obj_list) == 1 and self.obj_list[0].is_analytic_x

    def _lower(self, ctx, jac):
        # Convolution._drawKImage (galsim/convolve.py:420-426)
        self.obj_list[0]._lower(ctx, jac)
        for obj in self.obj_list[1:]:
            obj._lower(ctx, jac)
            ctx.emit(_lib.OP_MUL)

    def __repr__(self):
        return 'galsim_amd.Convolution(%r)' % (self.obj_list,)
