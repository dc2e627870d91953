"""Convolution (galsim/convolve.py), Fourier-space convolution only."""
import warnings

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
        if hard_edge and len(args) == 2:
            # galsim/convolve.py:171-175
            warnings.warn("Doing convolution of 2 objects, both with hard edges. "
                          "This might be more accurate and/or faster using real_space=True")
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


def Deconvolve(obj, gsparams=None, propagate_gsparams=True):
    """galsim/convolve.py:429-461 for a GSObject."""
    if not isinstance(obj, GSObject):
        raise TypeError("Argument to Deconvolve must be a GSObject.")
    return Deconvolution(obj, gsparams=gsparams, propagate_gsparams=propagate_gsparams)


class Deconvolution(GSObject):
    """galsim/convolve.py:464-612: the k-image is 1 / F(k) of the original where
    |F| > flux * kvalue_accuracy (else the inverse of that threshold), and 0 beyond the original's
    maxk.  Lowered as the original's program, an INV op and a KCLIP leaf (maxk cut)."""

    def __init__(self, obj, gsparams=None, propagate_gsparams=True):
        if not isinstance(obj, GSObject):
            raise TypeError("Argument to Deconvolution must be a GSObject.")
        self._gsparams = GSParams.check(gsparams, obj.gsparams)
        self._min_acc_kvalue = obj.flux * self.gsparams.kvalue_accuracy
        self._inv_min_acc_kvalue = 1. / self._min_acc_kvalue
        self._propagate_gsparams = propagate_gsparams
        self._orig_obj = obj.withGSParams(self._gsparams) if propagate_gsparams else obj

    @property
    def orig_obj(self):
        return self._orig_obj

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        if self._propagate_gsparams:
            ret._orig_obj = self._orig_obj.withGSParams(ret._gsparams)
        return ret

    @property
    def _maxk(self):
        return self.orig_obj.maxk

    @property
    def _stepk(self):
        return self.orig_obj.stepk

    @property
    def _flux(self):
        return 1. / self.orig_obj.flux

    def _lower(self, ctx, jac):
        from ._leaf import jac_record
        self.orig_obj._lower(ctx, jac)
        ctx.emit(_lib.OP_INV, 0, self._min_acc_kvalue, self._inv_min_acc_kvalue)
        if getattr(ctx, 'kvalue_mode', False):
            return                      # Deconvolution._kValue has no maxk cut
        if ctx.dk is None:
            raise _lib.GalSimAMDError("a Deconvolution needs the k-image spacing when lowered")
        from .render import LeafRecord
        rec = LeafRecord()
        sep, j = jac_record(jac)
        rec['type'] = _lib.LEAF_KCLIP
        rec['separable'] = sep
        rec['jac'] = j
        rec['cliprows'] = -1
        rec['scale'] = float(ctx.dk)
        rec['p'][0] = self.maxk ** 2
        rec['flux'] = 1.
        ctx.emit(_lib.OP_LEAF, ctx.add_leaf(rec))
        ctx.emit(_lib.OP_MUL)

    def __repr__(self):
        return 'galsim_amd.Deconvolution(%r)' % (self.orig_obj,)


def _auto_checks(obj, real_space, what):
    if not isinstance(obj, GSObject):
        raise TypeError("Argument to %s must be a GSObject." % what)
    if real_space is None:
        real_space = obj.has_hard_edges
    elif bool(real_space) != real_space:
        raise TypeError("real_space must be a boolean")
    if real_space and obj.is_analytic_x:
        raise NotImplementedError("real-space convolution is not part of galsim_amd")


def AutoConvolve(obj, real_space=None, gsparams=None, propagate_gsparams=True):
    """galsim/convolve.py:615-644 for a GSObject."""
    if not isinstance(obj, GSObject):
        raise TypeError("Argument to AutoConvolve must be a GSObject.")
    return AutoConvolution(obj, real_space, gsparams, propagate_gsparams)


class AutoConvolution(Convolution):
    """galsim/convolve.py:646-771: Convolution([obj, obj]) (the factor drawn twice and
    multiplied, as Convolution._drawKImage does)."""

    def __init__(self, obj, real_space=None, gsparams=None, propagate_gsparams=True):
        _auto_checks(obj, real_space, "AutoConvolution")
        self._real_space = False
        self._gsparams = GSParams.check(gsparams, obj.gsparams)
        self._propagate_gsparams = propagate_gsparams
        self._orig_obj = obj.withGSParams(self._gsparams) if propagate_gsparams else obj
        self._obj_list = [self._orig_obj, self._orig_obj]

    @property
    def orig_obj(self):
        return self._orig_obj

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        if self._propagate_gsparams:
            ret._orig_obj = self._orig_obj.withGSParams(ret._gsparams)
        ret._obj_list = [ret._orig_obj, ret._orig_obj]
        return ret

    def __repr__(self):
        return 'galsim_amd.AutoConvolution(%r)' % (self.orig_obj,)


def AutoCorrelate(obj, real_space=None, gsparams=None, propagate_gsparams=True):
    """galsim/convolve.py:774-803 for a GSObject."""
    if not isinstance(obj, GSObject):
        raise TypeError("Argument to AutoCorrelate must be a GSObject.")
    return AutoCorrelation(obj, real_space, gsparams, propagate_gsparams)


class AutoCorrelation(Convolution):
    """galsim/convolve.py:805-935: Convolution([obj, obj.transform(-1, 0, 0, -1)])."""

    def __init__(self, obj, real_space=None, gsparams=None, propagate_gsparams=True):
        _auto_checks(obj, real_space, "AutoCorrelation")
        self._real_space = False
        self._gsparams = GSParams.check(gsparams, obj.gsparams)
        self._propagate_gsparams = propagate_gsparams
        self._orig_obj = obj.withGSParams(self._gsparams) if propagate_gsparams else obj
        self._obj_list = [self._orig_obj, self._orig_obj.transform(-1, 0, 0, -1)]

    @property
    def orig_obj(self):
        return self._orig_obj

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        if self._propagate_gsparams:
            ret._orig_obj = self._orig_obj.withGSParams(ret._gsparams)
        ret._obj_list = [ret._orig_obj, ret._orig_obj.transform(-1, 0, 0, -1)]
        return ret

    def __repr__(self):
        return 'galsim_amd.AutoCorrelation(%r)' % (self.orig_obj,)
