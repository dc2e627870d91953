"""Sum (galsim/sum.py)."""
import numpy as np

from . import _lib
from .gsobject import GSObject
from .gsparams import GSParams


def Add(*args, **kwargs):
    if len(args) == 0:
        raise TypeError("At least one GSObject must be provided.")
    if len(args) == 1 and isinstance(args[0], (list, tuple)):
        args = args[0]
    return Sum(*args, **kwargs)


class Sum(GSObject):
    """galsim/sum.py:81-347: the k-image is the sum of the components' k-images."""

    def __init__(self, *args, **kwargs):
        if len(args) == 1 and isinstance(args[0], (list, tuple)):
            args = args[0]
        if len(args) == 0:
            raise TypeError("At least one GSObject must be provided.")
        gsparams = kwargs.pop("gsparams", None)
        self._propagate_gsparams = kwargs.pop('propagate_gsparams', True)
        if kwargs:
            raise TypeError("Sum got unexpected keyword argument(s): %s" % list(kwargs))
        for obj in args:
            if not isinstance(obj, GSObject):
                raise TypeError("Arguments to Sum must be GSObjects, not %s" % obj)
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
        return float(np.max([obj.maxk for obj in self.obj_list]))

    @property
    def _stepk(self):
        return float(np.min([obj.stepk for obj in self.obj_list]))

    @property
    def _flux(self):
        return float(np.sum([obj.flux for obj in self.obj_list]))

    @property
    def _is_analytic_x(self):
        return all(obj.is_analytic_x for obj in self.obj_list)

    @property
    def _has_hard_edges(self):
        return any(obj.has_hard_edges for obj in self.obj_list)

    def _lower(self, ctx, jac):
        # Sum._drawKImage (galsim/sum.py:341-347)
        self.obj_list[0]._lower(ctx, jac)
        for obj in self.obj_list[1:]:
            obj._lower(ctx, jac)
            ctx.emit(_lib.OP_ADD)

    def __repr__(self):
        return 'galsim_amd.Sum(%r)' % (self.obj_list,)
