"""Config-driven stamp batches.

GalSim's config pipeline builds one profile per object from the ``gal`` and ``psf`` fields
(galsim/config/gsobject.py: BuildGSObject, TransformObject), convolves them and draws each stamp
with ``DrawBasic`` -> ``drawImage`` (galsim/config/stamp.py:469-558).  Here the same fields are
evaluated for every object at once, as arrays, and lowered with vectorised numpy straight into
the stamp records and device program of one batched launch (``build_batch``): no per-object
Python objects are made, so planning thousands of stamps costs milliseconds.  ``build_profiles``
makes the per-object galsim_amd profiles from the same values, as the reference's config would;
the two plannings give the same records.

Supported (the FFT drawing path):

* objects: Exponential, Sersic (untruncated), Moffat (untruncated), Kolmogorov, Gaussian, Box,
  Pixel, Sum/Add, Convolution/Convolve, each with ``flux`` and the transformations dilate,
  ellip, rotate, scale_flux, shear, magnify and shift, applied in TransformObject's order;
* values: numbers, per-object arrays, and the value types List, Sequence, Random,
  RandomGaussian, XY (positions), G1G2 / E1E2 / GBeta (shears).  Random values come from
  numpy's generator seeded by ``seed``: the distributions of GalSim's config, not its random
  stream;
* ``image.pixel_scale``, ``stamp.size`` (or ``xsize`` / ``ysize``), ``stamp.offset`` (or
  ``image.offset``), ``stamp.draw_method`` ('auto' or 'fft').
"""
import math

import numpy as np

from . import _lib
from . import render
from .batch import (STAMP_BAD_PARAMS, STAMP_FFT_TOO_LARGE, BatchPlan, HostPlan,
                    _nonfinite_stamps)
from .gsparams import GSParams

_GSP = GSParams()


# -- values (galsim/config/value.py) ----------------------------------------------------------
class _Values:
    """Evaluates config values for objects 0..n-1."""

    def __init__(self, n, seed):
        self.n = n
        self.rng = np.random.default_rng(seed)

    def num(self, spec):
        n = self.n
        if isinstance(spec, (int, float, np.integer, np.floating)):
            return np.full(n, float(spec))
        if isinstance(spec, (list, tuple, np.ndarray)):
            a = np.asarray(spec, dtype=float)
            if a.shape != (n,):
                raise ValueError("a per-object value needs %d entries" % n)
            return a.copy()
        if not isinstance(spec, dict) or 'type' not in spec:
            raise ValueError("cannot evaluate config value %r" % (spec,))
        t = spec['type']
        if t == 'List':
            items = spec['items']
            idx = (self.index(spec.get('index')) % len(items)).astype(np.int64)
            vals = np.stack([self.num(v) for v in items])
            return vals[idx, np.arange(n)]
        if t == 'Sequence':
            return self.index(spec).astype(float)
        if t == 'Random':
            return self.rng.uniform(float(spec['min']), float(spec['max']), n)
        if t == 'RandomGaussian':
            return self.rng.normal(float(spec.get('mean', 0.)), float(spec['sigma']), n)
        raise ValueError("value type %r is not supported here" % t)

    def index(self, spec):
        """Sequence values (first + step * (k // repeat)); the default index of a List."""
        spec = spec or {'type': 'Sequence'}
        if not isinstance(spec, dict) or spec.get('type') != 'Sequence':
            return self.num(spec)
        k = np.arange(self.n) // int(spec.get('repeat', 1))
        return float(spec.get('first', 0)) + float(spec.get('step', 1)) * k

    def angle(self, spec):
        """Radians; {'type': 'Random'} without bounds is uniform on [0, 2 pi)."""
        if isinstance(spec, dict) and spec.get('type') == 'Random' and 'min' not in spec:
            return self.rng.uniform(0., 2. * math.pi, self.n)
        return self.num(spec)

    def position(self, spec):
        if isinstance(spec, dict) and spec.get('type') == 'XY':
            return self.num(spec.get('x', 0.)), self.num(spec.get('y', 0.))
        a = np.asarray(spec, dtype=float)
        if a.shape == (2,):
            return np.full(self.n, a[0]), np.full(self.n, a[1])
        if a.shape == (self.n, 2):
            return a[:, 0].copy(), a[:, 1].copy()
        raise ValueError("cannot evaluate position %r" % (spec,))

    def shear(self, spec):
        """(g1, g2) arrays of a shear value (galsim/config/value.py G1G2, E1E2, GBeta)."""
        t = spec.get('type')
        if t == 'G1G2':
            return self.num(spec.get('g1', 0.)), self.num(spec.get('g2', 0.))
        if t == 'GBeta':
            g, beta = self.num(spec['g']), self.angle(spec['beta'])
            return g * np.cos(2. * beta), g * np.sin(2. * beta)
        if t == 'E1E2':
            e1, e2 = self.num(spec.get('e1', 0.)), self.num(spec.get('e2', 0.))
            absesq = e1 ** 2 + e2 ** 2
            # Shear._e2g (galsim/shear.py:324-330)
            f = np.where(absesq > 1.e-4, 1. / (1. + np.sqrt(np.maximum(1. - absesq, 0.))),
                         0.5 + absesq * (0.125 + absesq * (0.0625 + absesq * 0.0390625)))
            return e1 * f, e2 * f
        raise ValueError("shear type %r is not supported here" % t)


# -- vectorised objects -----------------------------------------------------------------------
def _shear_matrix(g1, g2):
    """Shear.getMatrix (galsim/shear.py:365-386) for each object: (n, 2, 2)."""
    g = np.abs(g1 + 1j * g2)
    s = np.sqrt(1. - g ** 2)
    return np.stack([np.stack([(1. + g1) / s, g2 / s], -1),
                     np.stack([g2 / s, (1. - g1) / s], -1)], -2)


def _matmul(a, b):
    """Elementwise 2x2 products a . b."""
    out = np.empty(np.broadcast_shapes(a.shape, b.shape))
    out[..., 0, 0] = a[..., 0, 0] * b[..., 0, 0] + a[..., 0, 1] * b[..., 1, 0]
    out[..., 0, 1] = a[..., 0, 0] * b[..., 0, 1] + a[..., 0, 1] * b[..., 1, 1]
    out[..., 1, 0] = a[..., 1, 0] * b[..., 0, 0] + a[..., 1, 1] * b[..., 1, 0]
    out[..., 1, 1] = a[..., 1, 0] * b[..., 0, 1] + a[..., 1, 1] * b[..., 1, 1]
    return out


class _VObj:
    """An object of the batch: per-object parameter arrays; ``take`` selects objects."""

    def take(self, idx):
        raise NotImplementedError


class _VLeaf(_VObj):
    fields = ()

    def take(self, idx):
        o = object.__new__(type(self))
        o.__dict__.update({k: (v[idx] if isinstance(v, np.ndarray) else v)
                           for k, v in self.__dict__.items()})
        return o

    def sig(self, jac):
        """Structure bits: a Box's separability decides whether its factor is hoisted."""
        return []


def _jac_sep(jac, n):
    """(separable, (n, 4) jac) of jac_record (galsim_amd/_leaf.py) per object."""
    if jac is None:
        return np.ones(n, dtype=np.int32), np.tile([1., 0., 0., 1.], (n, 1))
    sep = ((jac[:, 0, 1] == 0.) & (jac[:, 1, 0] == 0.)).astype(np.int32)
    j = jac.reshape(-1, 4).copy()
    j[sep == 1, 1] = 0.
    j[sep == 1, 2] = 0.
    return sep, j


class VExponential(_VLeaf):
    def __init__(self, r0, flux):
        self.r0, self.flux = r0, flux

    def maxk(self):
        return (_GSP.maxk_threshold ** (-1. / 3.)) / self.r0

    def stepk(self):
        logx = math.log(_GSP.folding_threshold)
        R = -logx
        for _ in range(3):
            R = math.log(1. + R) - logx
        R = max(R, _GSP.stepk_minimum_hlr * 1.6783469900166605)
        return (math.pi / R) / self.r0

    def leaf(self, vc, jac):
        n = len(self.r0)
        sep, j = _jac_sep(jac, n)
        ksq_max = _GSP.kvalue_accuracy ** (-1. / 1.5) - 1.
        clip = np.full(n, -1, dtype=np.int32)
        if vc.precision == 'f64':
            for k in range(n):
                vc.tctx.Nk, vc.tctx.dk = int(vc.Nk[k]), float(vc.dk[k])
                jk = None if jac is None else jac[k]
                clip[k] = vc.tctx.exponential_cliprows(float(self.r0[k]), _GSP, jk)
        return dict(type=_lib.LEAF_EXPONENTIAL, separable=sep, jac=j, scale=self.r0,
                    flux=self.flux, cliprows=clip,
                    p=np.tile([ksq_max, math.sqrt(ksq_max), 0., 0., 0., 0.], (n, 1)))


class VSersic(_VLeaf):
    def __init__(self, n, r0, flux):
        self.n, self.r0, self.flux = n, r0, flux

    def _info(self, what):
        from .sersic import sersic_info
        out = np.empty(len(self.n))
        for nv in np.unique(self.n):
            m = self.n == nv
            out[m] = float(sersic_info(float(nv), _GSP, 0.)[0][what])
        return out

    def maxk(self):
        return self._info(0) / self.r0

    def stepk(self):
        return self._info(1) / self.r0

    def leaf(self, vc, jac):
        from .sersic import sersic_info
        nn = len(self.n)
        sep, j = _jac_sep(jac, nn)
        table = np.zeros(nn, dtype=np.int32)
        ntab = np.zeros(nn, dtype=np.int32)
        clip = np.full(nn, -1, dtype=np.int32)
        p = np.zeros((nn, 6))
        for nv in np.unique(self.n):
            m = self.n == nv
            sc, args, vals, y2 = sersic_info(float(nv), _GSP, 0.)
            key = (float(nv), 0., _GSP)
            table[m] = vc.tctx.add_table(key, args, vals, y2)
            ntab[m] = len(args)
            p[m] = [sc[4], sc[5], sc[6], sc[7], sc[8], sc[9]]
            if vc.precision == 'f32' and len(args) > 1:
                clip[m] = vc.tctx.add_spline_coefs(key, args, vals, y2)   # relative: see lower
        return dict(type=_lib.LEAF_SERSIC, separable=sep, jac=j, scale=self.r0, flux=self.flux,
                    table=table, ntab=ntab, cliprows=clip, p=p, _coef=vc.precision == 'f32')


class VMoffat(_VLeaf):
    def __init__(self, beta, r0, flux):
        self.beta, self.r0, self.flux = beta, r0, flux
        self._info = None

    def take(self, idx):
        o = _VLeaf.take(self, idx)
        o._info = None if self._info is None else self._info[idx]
        return o

    def info(self):
        if self._info is None:
            g = _lib.ptr(_lib.gsp_array(_GSP))
            out = np.zeros((len(self.beta), 5))
            for k in range(len(self.beta)):
                _lib.check(_lib.lib().gsmi_moffat_info(float(self.beta[k]), float(self.r0[k]),
                                                       float(self.flux[k]), g,
                                                       _lib.ptr(out[k])), "Moffat info")
            self._info = out
        return self._info

    def maxk(self):
        return self.info()[:, 0].copy()

    def stepk(self):
        return self.info()[:, 1].copy()

    def leaf(self, vc, jac):
        from .moffat import temme_constants
        n = len(self.beta)
        sep, j = _jac_sep(jac, n)
        info = self.info()
        kind = info[:, 4].astype(np.int32)
        p = np.zeros((n, 6))
        p[:, 0] = self.beta
        p[:, 1] = np.where(info[:, 2] != 0, self.flux / np.where(info[:, 2] != 0, info[:, 2], 1.),
                           0.)
        for k in np.nonzero(kind == 0)[0]:
            p[k, 2:6] = temme_constants(float(self.beta[k]) - 1.)
        return dict(type=_lib.LEAF_MOFFAT, separable=sep, jac=j, scale=self.r0,
                    flux=info[:, 2].copy(), kind=kind, p=p)


class VKolmogorov(_VLeaf):
    def __init__(self, lor0, flux):
        self.lor0, self.flux = lor0, flux

    def k0(self):
        return 2.992939911888651 / self.lor0

    def maxk(self):
        return (-math.log(_GSP.kvalue_accuracy)) ** 0.6 * self.k0()

    def stepk(self):
        return float(_kolmogorov_stepk_unit()) * self.k0()

    def leaf(self, vc, jac):
        n = len(self.lor0)
        sep, j = _jac_sep(jac, n)
        return dict(type=_lib.LEAF_KOLMOGOROV, separable=sep, jac=j, scale=1. / self.k0(),
                    flux=self.flux, p=np.zeros((n, 6)))


def _kolmogorov_stepk_unit():
    """KolmogorovInfo's stepk in units of k0 (galsim_amd.kolmogorov's cached value)."""
    from .kolmogorov import Kolmogorov, _KOLM_CACHE
    Kolmogorov(lam_over_r0=1.)._stepk
    return _KOLM_CACHE[_GSP][1]


class VGaussian(_VLeaf):
    def __init__(self, sigma, flux):
        self.sigma, self.flux = sigma, flux

    def maxk(self):
        return math.sqrt(-2. * math.log(_GSP.maxk_threshold)) / self.sigma

    def stepk(self):
        R = max(math.sqrt(-2. * math.log(_GSP.folding_threshold)),
                _GSP.stepk_minimum_hlr * 1.1774100225154747)
        return math.pi / (R * self.sigma)

    def leaf(self, vc, jac):
        n = len(self.sigma)
        sep, j = _jac_sep(jac, n)
        kva = _GSP.kvalue_accuracy
        return dict(type=_lib.LEAF_GAUSSIAN, separable=sep, jac=j, scale=self.sigma,
                    flux=self.flux,
                    p=np.tile([-2. * math.log(kva), (kva * 48.) ** (1. / 3.), 0., 0., 0., 0.],
                              (n, 1)))


class VBox(_VLeaf):
    def __init__(self, w, h, flux):
        self.w, self.h, self.flux = w, h, flux

    def maxk(self):
        return 2. / (_GSP.maxk_threshold * np.minimum(self.w, self.h))

    def stepk(self):
        return math.pi / np.maximum(self.w, self.h)

    def sig(self, jac):
        n = len(self.w)
        return [_jac_sep(jac, n)[0].astype(bool)]

    def leaf(self, vc, jac):
        n = len(self.w)
        sep, j = _jac_sep(jac, n)
        return dict(type=_lib.LEAF_BOX, separable=sep, jac=j, scale=self.w / (2. * math.pi),
                    scale2=self.h / (2. * math.pi), flux=self.flux, p=np.zeros((n, 6)))


class VTransform(_VObj):
    """Transformation (galsim_amd/transform.py) for every object, merged with an inner
    transformation as the constructor does."""

    def __init__(self, obj, jac=None, dx=None, dy=None, flux_ratio=None, n=None):
        n = n if n is not None else _count(obj)
        self.jac = jac
        self.dx = np.zeros(n) if dx is None else dx.copy()
        self.dy = np.zeros(n) if dy is None else dy.copy()
        self.flux_ratio = np.ones(n) if flux_ratio is None else flux_ratio.copy()
        if isinstance(obj, VTransform):
            has = (obj.dx != 0.) | (obj.dy != 0.)
            if jac is None:
                dx1, dy1 = obj.dx, obj.dy
            else:
                # _fwd_normal (galsim/transform.py:436-444)
                temp = jac[:, 0, 1] * obj.dy
                dy1 = obj.dy * jac[:, 1, 1] + jac[:, 1, 0] * obj.dx
                dx1 = obj.dx * jac[:, 0, 0] + temp
            self.dx = np.where(has, self.dx + dx1, self.dx)
            self.dy = np.where(has, self.dy + dy1, self.dy)
            if jac is None:
                self.jac = obj.jac
            elif obj.jac is not None:
                self.jac = _matmul(jac, obj.jac)
            self.flux_ratio = self.flux_ratio * obj.flux_ratio
            self.orig = obj.orig
        else:
            self.orig = obj

    def take(self, idx):
        o = object.__new__(VTransform)
        o.jac = None if self.jac is None else self.jac[idx]
        o.dx, o.dy, o.flux_ratio = self.dx[idx], self.dy[idx], self.flux_ratio[idx]
        o.orig = self.orig.take(idx)
        return o

    def det(self):
        j = self.jac
        if j is None:
            return np.ones(len(self.dx))
        diag = (j[:, 0, 1] == 0.) & (j[:, 1, 0] == 0.)
        ident = diag & (j[:, 0, 0] == 1.) & (j[:, 1, 1] == 1.)
        return np.where(ident, 1., np.where(diag, j[:, 0, 0] * j[:, 1, 1],
                                            j[:, 0, 0] * j[:, 1, 1] - j[:, 0, 1] * j[:, 1, 0]))

    def flux_scaling(self):
        return np.abs(self.det()) * self.flux_ratio

    def flux(self):
        return self.flux_scaling() * _flux(self.orig)

    def _major_minor(self):
        j = self.jac
        if j is None:
            return 1., 1.
        h1 = np.hypot(j[:, 0, 0] + j[:, 1, 1], j[:, 0, 1] - j[:, 1, 0])
        h2 = np.hypot(j[:, 0, 0] - j[:, 1, 1], j[:, 0, 1] + j[:, 1, 0])
        return 0.5 * np.abs(h1 + h2), 0.5 * np.abs(h1 - h2)

    def maxk(self):
        return _maxk(self.orig) / self._major_minor()[1]

    def stepk(self):
        stepk = _stepk(self.orig) / self._major_minor()[0]
        has = (self.dx != 0.) | (self.dy != 0.)
        dr = np.hypot(self.dx, self.dy)
        return np.where(has, math.pi / (math.pi / stepk + dr), stepk)

    def _outer(self, jac):
        if jac is None:
            return self.jac
        if self.jac is None:
            return jac
        return _matmul(jac, self.jac)

    def sig(self, jac):
        fs = self.flux_scaling()
        return (_sig(self.orig, self._outer(jac)) +
                [(self.dx != 0.) | (self.dy != 0.),
                 np.abs(fs - 1.) > _GSP.kvalue_accuracy])


class _VComposite(_VObj):
    def __init__(self, items):
        self.items = items

    def take(self, idx):
        o = object.__new__(type(self))
        o.items = [it.take(idx) for it in self.items]
        return o

    def sig(self, jac):
        return [b for it in self.items for b in _sig(it, jac)]


class VConvolve(_VComposite):
    def maxk(self):
        return np.minimum.reduce([_maxk(it) for it in self.items])

    def stepk(self):
        acc = 0
        for it in self.items:
            acc = acc + _stepk(it) ** (-2)
        return acc ** (-0.5)

    def flux(self):
        f = _flux(self.items[0])
        for it in self.items[1:]:
            f = f * _flux(it)
        return f


class VSum(_VComposite):
    def maxk(self):
        return np.maximum.reduce([_maxk(it) for it in self.items])

    def stepk(self):
        return np.minimum.reduce([_stepk(it) for it in self.items])

    def flux(self):
        f = _flux(self.items[0])
        for it in self.items[1:]:
            f = f + _flux(it)
        return f


def _maxk(o):
    return np.asarray(o.maxk(), dtype=float)


def _stepk(o):
    return np.asarray(o.stepk(), dtype=float)


def _flux(o):
    return o.flux() if callable(o.flux) else o.flux


def _sig(o, jac):
    return o.sig(jac)


def _count(o):
    if isinstance(o, VTransform):
        return len(o.dx)
    if isinstance(o, _VComposite):
        return _count(o.items[0])
    for v in o.__dict__.values():
        if isinstance(v, np.ndarray):
            return len(v)
    raise ValueError("empty object")


# -- lowering of a group of objects with one program structure ---------------------------------
class _VLower:
    """The vectorised counterpart of render.Lowering for m objects whose programs have one
    structure: ops as (code, arg, a, b, c) with per-object constant arrays, leaves as dicts of
    per-object field arrays."""

    def __init__(self, precision, tctx, Nk, dk):
        self.precision = precision
        self.tctx = tctx           # tables, spline coefficients and clip rows (render.Lowering)
        self.Nk, self.dk = Nk, dk
        self.ops = []
        self.leaves = []

    def emit(self, code, arg=0, a=0., b=0., c=0.):
        self.ops.append((code, arg, a, b, c))

    def lower(self, o, jac):
        if isinstance(o, VTransform):
            # Transformation._lower (galsim_amd/transform.py)
            self.lower(o.orig, o._outer(jac))
            fs = o.flux_scaling()
            has = (o.dx != 0.) | (o.dy != 0.)
            if has[0]:
                if jac is None:
                    ux, uy = o.dx, o.dy
                else:
                    ux = jac[:, 0, 0] * o.dx + jac[:, 0, 1] * o.dy
                    uy = jac[:, 1, 0] * o.dx + jac[:, 1, 1] * o.dy
                self.emit(_lib.OP_PHASE, 0, ux, uy, fs)
            elif abs(fs[0] - 1.) > _GSP.kvalue_accuracy:
                self.emit(_lib.OP_SCALE, 0, fs)
        elif isinstance(o, VConvolve):
            self.lower(o.items[0], jac)
            for it in o.items[1:]:
                self.lower(it, jac)
                self.emit(_lib.OP_MUL)
        elif isinstance(o, VSum):
            self.lower(o.items[0], jac)
            for it in o.items[1:]:
                self.lower(it, jac)
                self.emit(_lib.OP_ADD)
        else:
            self.leaves.append(o.leaf(self, jac))
            self.emit(_lib.OP_LEAF, len(self.leaves) - 1)

    def hoist(self):
        """render.hoist_axis_factors on the shared structure: (op order, op_count, axis)."""
        prog = list(range(len(self.ops)))
        axis = []
        while True:
            if len(prog) > 1 and self.ops[prog[-1]][0] in (_lib.OP_PHASE, _lib.OP_SCALE):
                axis.append(prog.pop())
                continue
            if (len(prog) > 2 and self.ops[prog[-1]][0] == _lib.OP_MUL and
                    self.ops[prog[-2]][0] == _lib.OP_LEAF):
                lf = self.leaves[self.ops[prog[-2]][1]]
                if lf['type'] == _lib.LEAF_BOX and bool(np.all(lf['separable'])):
                    axis.append(prog[-2])
                    del prog[-2:]
                    continue
            break
        return prog + axis, len(prog), len(axis)


# -- spec -> objects ---------------------------------------------------------------------------
_TRANSFORMS = ('dilate', 'dilation', 'ellip', 'rotate', 'rotation', 'scale_flux', 'shear',
               'magnify', 'magnification', 'shift')


def _need(spec, *names):
    have = [k for k in names if k in spec]
    if len(have) != 1:
        raise ValueError("%s needs exactly one of %s" % (spec.get('type'), names))
    return have[0]


def _eval_obj(spec, V):
    """(vector object, per-object builder f(i) -> galsim_amd object) of a config object."""
    import galsim_amd as g
    if 'gsparams' in spec:
        raise ValueError("config gsparams are not supported by the batched builder")
    t = spec['type']
    flux_given = 'flux' in spec
    flux = V.num(spec.get('flux', 1.))
    if t == 'Exponential':
        k = _need(spec, 'half_light_radius', 'scale_radius')
        v = V.num(spec[k])
        r0 = v / 1.6783469900166605 if k == 'half_light_radius' else v
        obj = VExponential(r0, flux)
        mk = lambda i: g.Exponential(scale_radius=float(r0[i]), flux=float(flux[i]))
    elif t == 'Sersic':
        if spec.get('trunc', 0.) not in (0, 0.):
            raise ValueError("a truncated Sersic is not supported by the batched builder")
        nn = V.num(spec['n'])
        k = _need(spec, 'half_light_radius', 'scale_radius')
        v = V.num(spec[k])
        if k == 'half_light_radius':
            fac = np.empty(V.n)
            for nv in np.unique(nn):
                out = np.zeros(1)
                _lib.check(_lib.lib().gsmi_sersic_hlr(float(nv), 1., _lib.ptr(out)),
                           "SersicHLR")
                fac[nn == nv] = out[0]
            r0 = v / fac
        else:
            r0 = v
        obj = VSersic(nn, r0, flux)
        mk = lambda i: g.Sersic(float(nn[i]), scale_radius=float(r0[i]), flux=float(flux[i]))
    elif t == 'Moffat':
        if spec.get('trunc', 0.) not in (0, 0.):
            raise ValueError("a truncated Moffat is not supported by the batched builder")
        beta = V.num(spec['beta'])
        k = _need(spec, 'fwhm', 'half_light_radius', 'scale_radius')
        v = V.num(spec[k])
        if k == 'fwhm':
            r0 = v / (2. * np.sqrt(2. ** (1. / beta) - 1.))
        elif k == 'scale_radius':
            r0 = v
        else:
            r0 = np.empty(V.n)
            out = np.zeros(1)
            for i in range(V.n):
                _lib.check(_lib.lib().gsmi_moffat_scale_radius_from_hlr(
                    float(v[i]), float(beta[i]), _lib.ptr(out)), "Moffat hlr")
                r0[i] = out[0]
        obj = VMoffat(beta, r0, flux)
        mk = lambda i: g.Moffat(float(beta[i]), scale_radius=float(r0[i]), flux=float(flux[i]))
    elif t == 'Kolmogorov':
        k = _need(spec, 'fwhm', 'lam_over_r0', 'half_light_radius')
        v = V.num(spec[k])
        lor0 = (v / 0.9758634299 if k == 'fwhm' else v / 0.5548101137
                if k == 'half_light_radius' else v)
        obj = VKolmogorov(lor0, flux)
        mk = lambda i: g.Kolmogorov(lam_over_r0=float(lor0[i]), flux=float(flux[i]))
    elif t == 'Gaussian':
        k = _need(spec, 'sigma', 'fwhm', 'half_light_radius')
        v = V.num(spec[k])
        sigma = (v / 2.3548200450309493 if k == 'fwhm' else v / 1.1774100225154747
                 if k == 'half_light_radius' else v)
        obj = VGaussian(sigma, flux)
        mk = lambda i: g.Gaussian(sigma=float(sigma[i]), flux=float(flux[i]))
    elif t in ('Box', 'Pixel'):
        if t == 'Pixel':
            w = h = V.num(spec['scale'])
        else:
            w, h = V.num(spec['width']), V.num(spec['height'])
        obj = VBox(w, h, flux)
        mk = lambda i: g.Box(float(w[i]), float(h[i]), flux=float(flux[i]))
    elif t in ('Sum', 'Add', 'Convolution', 'Convolve'):
        subs = [_eval_obj(it, V) for it in spec['items']]
        if len(subs) == 1:
            obj, mk = subs[0]
        else:
            cls = VSum if t in ('Sum', 'Add') else VConvolve
            obj = cls([s[0] for s in subs])
            if t in ('Sum', 'Add'):
                mk = lambda i: g.Add([s[1](i) for s in subs])
            else:
                mk = lambda i: g.Convolve([s[1](i) for s in subs])
        if flux_given:
            # withFlux (galsim_amd/gsobject.py): a flux ratio, none when it is exactly 1
            ratio = flux / _flux(obj)
            obj0, mk0 = obj, mk
            obj = VTransform(obj0, flux_ratio=ratio, n=V.n)
            mk = lambda i: mk0(i).withFlux(float(flux[i]))
    else:
        raise ValueError("object type %r is not supported by the batched builder" % t)
    return _apply_transforms(spec, obj, mk, V)


def _apply_transforms(spec, obj, mk, V):
    """TransformObject's order (galsim/config/gsobject.py:515-549)."""
    n = V.n
    for key in _TRANSFORMS:
        if key not in spec:
            continue
        val = spec[key]
        obj0, mk0 = obj, mk
        if key in ('dilate', 'dilation'):
            s = V.num(val)
            jac = np.zeros((n, 2, 2))
            jac[:, 0, 0] = jac[:, 1, 1] = s
            obj = VTransform(obj0, jac=jac, flux_ratio=s ** -2., n=n)
            mk = (lambda mk0, s: lambda i: mk0(i).dilate(float(s[i])))(mk0, s)
        elif key in ('ellip', 'shear'):
            g1, g2 = V.shear(val)
            obj = VTransform(obj0, jac=_shear_matrix(g1, g2), n=n)
            mk = (lambda mk0, g1, g2: lambda i: mk0(i).shear(g1=float(g1[i]),
                                                             g2=float(g2[i])))(mk0, g1, g2)
        elif key in ('rotate', 'rotation'):
            th = V.angle(val)
            s, c = np.sin(th), np.cos(th)
            jac = np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)
            obj = VTransform(obj0, jac=jac, n=n)
            mk = (lambda mk0, th: lambda i: mk0(i).rotate(float(th[i])))(mk0, th)
        elif key == 'scale_flux':
            r = V.num(val)
            obj = VTransform(obj0, flux_ratio=r, n=n)
            mk = (lambda mk0, r: lambda i: mk0(i) * float(r[i]))(mk0, r)
        elif key in ('magnify', 'magnification'):
            mu = V.num(val)
            s = np.sqrt(mu)
            jac = np.zeros((n, 2, 2))
            jac[:, 0, 0] = jac[:, 1, 1] = s
            obj = VTransform(obj0, jac=jac, n=n)
            mk = (lambda mk0, mu: lambda i: mk0(i).magnify(float(mu[i])))(mk0, mu)
        elif key == 'shift':
            dx, dy = V.position(val)
            obj = VTransform(obj0, dx=dx, dy=dy, n=n)
            mk = (lambda mk0, dx, dy: lambda i: mk0(i)._shift(float(dx[i]),
                                                              float(dy[i])))(mk0, dx, dy)
    return obj, mk


# -- the batch ---------------------------------------------------------------------------------
def _evaluate(config, nobjects, seed):
    V = _Values(int(nobjects), seed)
    parts = [_eval_obj(config[k], V) for k in ('gal', 'psf') if k in config]
    if not parts:
        raise ValueError("the config needs a gal or psf field")
    if len(parts) == 1:
        prof, mk = parts[0]
    else:
        prof = VConvolve([p[0] for p in parts])
        mk = lambda i: __import__('galsim_amd').Convolve([p[1](i) for p in parts])
    image = config.get('image', {})
    stamp = config.get('stamp', {})
    scale = float(image.get('pixel_scale', 1.))
    if 'size' in stamp:
        nx = ny = V.num(stamp['size'])
    else:
        nx, ny = V.num(stamp['xsize']), V.num(stamp['ysize'])
    nx, ny = nx.astype(np.int64), ny.astype(np.int64)
    off = stamp.get('offset', image.get('offset'))
    ox, oy = (np.zeros(V.n), np.zeros(V.n)) if off is None else V.position(off)
    method = stamp.get('draw_method', 'auto')
    if method not in ('auto', 'fft'):
        raise ValueError("draw_method %r is not batched here" % method)
    return prof, mk, scale, nx, ny, ox, oy, method


def build_profiles(config, nobjects, seed=0):
    """The per-object galsim_amd profiles and drawing arguments of a config:
    (profiles, nx, ny, scale, offsets, method)."""
    prof, mk, scale, nx, ny, ox, oy, method = _evaluate(config, nobjects, seed)
    return ([mk(i) for i in range(int(nobjects))], nx, ny, scale, np.stack([ox, oy], 1),
            method)


def plan_config(config, nobjects, precision='f32', seed=0):
    """The HostPlan of a config's stamps, lowered with vectorised numpy (galsim_amd.batch's
    per-object planner gives the same records)."""
    prof, mk, scale, nx, ny, ox, oy, method = _evaluate(config, nobjects, seed)
    n = int(nobjects)
    # the image-coordinate profile (drawImage, galsim/gsobject.py:1720-1777): offset
    # convention, pixel-scale transform, pixel
    dx = ox - np.where(nx % 2 == 0, 0.5, 0.)
    dy = oy - np.where(ny % 2 == 0, 0.5, 0.)
    jac = None
    if scale != 1.:
        jac = np.zeros((n, 2, 2))
        jac[:, 0, 0] = jac[:, 1, 1] = 1. / scale
    p = VTransform(prof, jac=jac, dx=dx, dy=dy, flux_ratio=np.full(n, scale ** 2), n=n)
    p = VConvolve([p, VBox(np.ones(n), np.ones(n), np.ones(n))])
    # centred bounds (BoundsI(1, nx, 1, ny) shifted by -center)
    xmin = 1 - (1 + nx // 2)
    ymin = 1 - (1 + ny // 2)
    xmax, ymax = nx + xmin - 1, ny + ymin - 1
    # drawFFT_makeKImage sizes (render.plan_sizes)
    host_status = np.zeros(n, dtype=np.int32)
    with np.errstate(all='ignore'):
        stepk, maxk = _stepk(p), _maxk(p)
        Nd = 2. * math.pi / (1.0 * stepk)
        good = np.isfinite(Nd) & np.isfinite(maxk) & (maxk > 0)
        N = np.zeros(n, dtype=np.int64)
        N[good] = np.ceil(Nd[good] * (1. - 1.e-12)).astype(np.int64)
        N = 2 * ((N + 1) // 2)
        image_N = np.maximum(np.max(np.abs(np.stack([xmin, xmax, ymin, ymax])), 0) * 2,
                             np.maximum(nx, ny))
        N = np.maximum(N, image_N)
        for v in np.unique(N[good]):
            N[good & (N == v)] = _lib.lib().gsmi_good_fft_size(int(v))
        N = np.maximum(N, _GSP.minimum_fft_size)
        dk = 2. * math.pi / (N * 1.0)
        Nk = np.where(N * dk / 2 > maxk, N,
                      np.ceil(np.where(good, maxk / dk, 0.)).astype(np.int64) * 2)
    host_status[~good] |= STAMP_BAD_PARAMS
    host_status[good & (Nk > _GSP.maximum_fft_size)] |= STAMP_FFT_TOO_LARGE
    # groups of one program structure
    sig = _sig(p, None)
    code = np.zeros(n, dtype=np.int64)
    for b, bits in enumerate(sig):
        code |= np.asarray(bits, dtype=np.int64) << b
    tctx = render.Lowering(precision)
    recs = np.zeros(n, dtype=_lib.STAMP_DTYPE)
    op_blocks, leaf_blocks = [], []
    nop = nleaf = 0
    keep = good.copy()
    for c in np.unique(code[good]):
        idx = np.nonzero(good & (code == c))[0]
        vc = _VLower(precision, tctx, Nk[idx], dk[idx])
        vc.lower(p.take(idx), None)
        order, nops, naxis = vc.hoist()
        m, k, L = len(idx), len(vc.ops), len(vc.leaves)
        ops = np.zeros((m, k), dtype=_lib.OP_DTYPE)
        for col, o in enumerate(order):
            code_, arg, a, b, cc = vc.ops[o]
            ops['code'][:, col] = code_
            ops['arg'][:, col] = arg
            ops['a'][:, col], ops['b'][:, col], ops['c'][:, col] = a, b, cc
        leaves = np.zeros((m, L), dtype=_lib.LEAF_DTYPE)
        for col, lf in enumerate(vc.leaves):
            for name, v in lf.items():
                if name == '_coef':
                    continue
                if name == 'jac' or name == 'p':
                    leaves[name][:, col, :] = v
                else:
                    leaves[name][:, col] = v
            if lf.get('_coef'):
                # spline coefficient offsets relative to the region -> absolute in tables32
                leaves['cliprows'][:, col] = np.where(leaves['cliprows'][:, col] >= 0,
                                                      leaves['cliprows'][:, col] + (1 << 30), -1)
        recs['op_begin'][idx] = nop + np.arange(m) * k
        recs['leaf_begin'][idx] = nleaf + np.arange(m) * L
        recs['op_count'][idx] = nops
        recs['axis_count'][idx] = naxis
        op_blocks.append(ops.reshape(-1))
        leaf_blocks.append(leaves.reshape(-1))
        nop += m * k
        nleaf += m * L
    ops = np.concatenate(op_blocks) if op_blocks else np.zeros(0, dtype=_lib.OP_DTYPE)
    leaves = np.concatenate(leaf_blocks) if leaf_blocks else np.zeros(0, dtype=_lib.LEAF_DTYPE)
    marked = leaves['cliprows'] >= (1 << 30)
    leaves['cliprows'][marked] += tctx._coef_base() - (1 << 30)
    recs['N'], recs['Nk'], recs['nx'], recs['ny'] = N, Nk, nx, ny
    recs['x0'], recs['y0'], recs['dk'] = xmin, ymin, dk
    kept = np.nonzero(keep)[0]
    for i in kept[_nonfinite_stamps(recs[kept], ops, leaves)]:
        host_status[i] |= STAMP_BAD_PARAMS
        keep[i] = False
    return HostPlan(n=n, precision=precision, recs=recs, keep=keep, host_status=host_status,
                    ops=ops, leaves=leaves,
                    tables=np.concatenate(tctx.tables) if tctx.tables else np.zeros(1),
                    tables32=tctx.tables32_array(),
                    clip=np.concatenate(tctx.clip) if tctx.clip else np.zeros(2, dtype=np.int32),
                    nx=nx, ny=ny, scale=scale, offsets=np.stack([ox, oy], 1), method=method,
                    use_true_center=True, profile=mk)


def build_batch(config, nobjects, precision='f32', seed=0):
    """A BatchPlan of a config's stamps (plan_config, then the device set-up)."""
    return BatchPlan(None, None, None, None, lowered=plan_config(config, nobjects, precision,
                                                                 seed))
