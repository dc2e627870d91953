"""Real-space drawing: GSObject.drawReal and drawImage(method='no_pixel' / 'sb') of analytic
profiles, evaluated at the pixel centres on the GPU (gsmi_draw_real_f64).

The reference's chain is GSObject.drawReal -> _drawReal of Transformation / Sum / a one-object
Convolution (galsim/gsobject.py:1846-1898, galsim/transform.py:571-583, galsim/sum.py:304-310,
galsim/convolve.py:395-397), ending at each leaf's SBProfile::draw (src/SBProfile.cpp:304-381).
``components`` walks the same tree with the same (jac, offset, flux_scaling) bookkeeping and turns
every leaf into one gsmi_xcomp record: draw()'s offset / Jacobian set-up and the leaf's own
normalisation and radius scaling, done here on the host in the reference's order; the device
evaluates the records at every pixel.  Leaves covered: Gaussian, Exponential, Moffat, Sersic
(DeVaucouleurs), Kolmogorov, Airy, Spergel, Box / Pixel and TopHat.
"""
import math

import numpy as np

from . import _lib

XCOMP_DTYPE = np.dtype([('type', np.int32), ('form', np.int32), ('kind', np.int32),
                        ('izero', np.int32), ('jzero', np.int32), ('ic', np.int32),
                        ('jc', np.int32), ('i1', np.int32), ('i2', np.int32), ('j1', np.int32),
                        ('j2', np.int32), ('pad', np.int32), ('x0', np.float64),
                        ('dx', np.float64), ('dxy', np.float64), ('y0', np.float64),
                        ('dy', np.float64), ('dyx', np.float64), ('norm', np.float64),
                        ('amp', np.float64), ('p', np.float64, (2,))], align=True)
assert XCOMP_DTYPE.itemsize == 128
XCOMP_MAX = 16


def _clamp_ceil(v, lo, hi):
    if not v == v:
        return hi
    c = math.ceil(v) if math.isfinite(v) else (hi if v > 0 else lo)
    return int(min(max(c, lo), hi))


def _draw_setup(xmin, ymin, m, n, imscale, jac, xoff, yoff, flux_ratio):
    """SBProfile::draw (src/SBProfile.cpp:304-381): the fill's coordinate form and the amplitude
    ratio applied after it.  Returns (form, x0, dx, dxy, y0, dy, dyx, izero, jzero, amp)."""
    if xoff == 0.:
        x0 = xmin * imscale
        izero = -xmin if xmin < 0 else 0
    else:
        x0 = xmin * imscale - xoff
        izero = int(-x0 / imscale + 0.5)
        if abs(x0 + izero * imscale) > 1.e-10 or izero < 0 or izero >= m:
            izero = 0
    if yoff == 0.:
        y0 = ymin * imscale
        jzero = -ymin if ymin < 0 else 0
    else:
        y0 = ymin * imscale - yoff
        jzero = int(-y0 / imscale + 0.5)
        if abs(y0 + jzero * imscale) > 1.e-10 or jzero < 0 or jzero >= n:
            jzero = 0
    if jac is None:
        if imscale != 1.:
            flux_ratio *= imscale * imscale
        return 0, x0, imscale, 0., y0, imscale, 0., izero, jzero, flux_ratio
    mA, mB, mC, mD = float(jac[0, 0]), float(jac[0, 1]), float(jac[1, 0]), float(jac[1, 1])
    if mB == 0. and mC == 0.:
        xscale = 1. / mA
        yscale = 1. / mD
        dx = imscale * xscale
        dy = imscale * yscale
        flux_ratio *= abs(dx * dy)
        return 0, x0 * xscale, dx, 0., y0 * yscale, dy, 0., izero, jzero, flux_ratio
    det = mA * mD - mB * mC
    invdet = 1. / det
    new_x0 = (mD * x0 - mB * y0) * invdet
    new_y0 = (-mC * x0 + mA * y0) * invdet
    invdet *= imscale
    flux_ratio *= abs(invdet * imscale)
    return 1, new_x0, mD * invdet, -mB * invdet, new_y0, mA * invdet, -mC * invdet, 0, 0, \
        flux_ratio


def _moffat_kind(beta, xacc):
    for kind, b in ((1, 1.), (15, 1.5), (2, 2.), (25, 2.5), (3, 3.), (35, 3.5), (4, 4.)):
        if abs(beta - b) < xacc:
            return kind
    return 0


_KOLM_XTABLE = {}


def kolmogorov_xtable(gsparams):
    """KolmogorovInfo's x-space radial table (gsmi_kolmogorov_xtable) as one float64 array:
    knots, values, second derivatives (cached per GSParams)."""
    if gsparams not in _KOLM_XTABLE:
        g = _lib.gsp_array(gsparams)
        L = _lib.lib()
        n = L.gsmi_kolmogorov_xtable(_lib.ptr(g), None, None, None, 0)
        if n < 0:
            _lib.check(n, "gsmi_kolmogorov_xtable")
        t = np.zeros(3 * n)
        got = L.gsmi_kolmogorov_xtable(_lib.ptr(g), _lib.ptr(t[:n]), _lib.ptr(t[n:2 * n]),
                                       _lib.ptr(t[2 * n:]), n)
        if got != n:
            _lib.check(got if got < 0 else -3, "gsmi_kolmogorov_xtable")
        _KOLM_XTABLE[gsparams] = t
    return _KOLM_XTABLE[gsparams]


def _leaf_component(leaf, m, n, setup, tables):
    """The gsmi_xcomp record of one leaf drawn with draw()'s set-up ``setup``; a Kolmogorov
    leaf appends its radial table to the list ``tables``."""
    from .airy import Airy
    from .kolmogorov import Kolmogorov
    from .box import Box, TopHat
    from .exponential import Exponential
    from .gaussian import Gaussian
    from .moffat import Moffat
    from .sersic import Sersic
    from .spergel import Spergel
    form, x0, dx, dxy, y0, dy, dyx, izero, jzero, amp = setup
    rec = np.zeros((), dtype=XCOMP_DTYPE)
    rec['form'] = form
    rec['ic'] = rec['jc'] = -1
    rec['amp'] = amp
    radial = None
    if isinstance(leaf, Gaussian):
        # src/SBGaussian.cpp:55-76
        inv = 1. / leaf._sigma
        rec['type'] = _lib.LEAF_GAUSSIAN
        rec['norm'] = leaf._flux * (inv * inv) / (2. * math.pi)
        radial = inv
    elif isinstance(leaf, Exponential):
        # src/SBExponential.cpp:70-90
        inv = 1. / leaf._r0
        rec['type'] = _lib.LEAF_EXPONENTIAL
        rec['norm'] = (leaf._flux / (2. * math.pi)) * (inv * inv)
        radial = inv
    elif isinstance(leaf, Moffat):
        # src/SBMoffat.cpp:170-224
        beta, rD, trunc = leaf._beta, leaf._r0, leaf._trunc
        inv = 1. / rD
        if trunc > 0.:
            maxRrD = trunc * inv
            flux_factor = 1. - math.pow(1. + maxRrD * maxRrD, 1. - beta)
        else:
            flux_factor = 1.
            maxRrD = math.sqrt(math.pow(leaf.gsparams.xvalue_accuracy, 1. / (1. - beta)) - 1.)
        rec['type'] = _lib.LEAF_MOFFAT
        rec['norm'] = leaf._flux * (beta - 1.) / (math.pi * flux_factor * (rD * rD))
        rec['kind'] = _moffat_kind(beta, leaf.gsparams.xvalue_accuracy)
        rec['p'] = (beta, maxRrD * maxRrD)
        radial = inv
    elif isinstance(leaf, Sersic):
        # src/SBSersic.cpp:70-90, 354-360 (SersicInfo with trunc in units of r0)
        nn, r0 = leaf._n, leaf._r0
        inv = 1. / r0
        t = leaf._trunc / r0
        xnorm = 1. / (2. * math.pi * nn * math.gamma(2. * nn) * leaf._flux_fraction)
        rec['type'] = _lib.LEAF_SERSIC
        rec['norm'] = (leaf._flux * xnorm) * (inv * inv)
        rec['p'] = (0.5 / nn, t * t if t > 0. else 0.)
        radial = inv
    elif isinstance(leaf, Airy):
        # src/SBAiry.cpp:66-80 (D = 1 / lam_over_diam, xnorm = flux D^2), 88-108, 410-429
        D = 1. / leaf._lod
        rec['type'] = _lib.LEAF_AIRY
        rec['norm'] = leaf._flux * (D * D)
        rec['p'] = (leaf._obscuration, math.sqrt(8. * leaf.gsparams.xvalue_accuracy))
        radial = D
    elif isinstance(leaf, Kolmogorov):
        # src/SBKolmogorov.cpp:83-115 (k0 = K0_FACTOR / lam_over_r0, xnorm = flux k0^2)
        k0 = leaf._k0
        t = kolmogorov_xtable(leaf.gsparams)
        rec['type'] = _lib.LEAF_KOLMOGOROV
        rec['norm'] = leaf._flux * (k0 * k0)
        rec['i1'] = sum(len(a) for a in tables)
        rec['i2'] = len(t) // 3
        tables.append(t)
        radial = k0
    elif isinstance(leaf, Spergel):
        # src/SBSpergel.cpp:80-95 (xnorm = flux getXNorm / r0^2), 525-535, 644-651
        nu, r0 = leaf._nu, leaf._r0
        g1 = math.gamma(nu + 1.)
        rec['type'] = _lib.LEAF_SPERGEL
        rec['norm'] = (leaf._flux * (math.pow(2., -nu) / g1 / (2.0 * math.pi))) / (r0 * r0)
        rec['p'] = (nu, g1 / (2. * nu) * math.pow(2., nu) if nu > 0. else math.inf)
        radial = 1. / r0
    elif isinstance(leaf, Box):
        # src/SBBox.cpp:55-148
        rec['type'] = _lib.LEAF_BOX
        rec['norm'] = leaf._flux / (leaf._width * leaf._height)
        wo2, ho2 = 0.5 * leaf._width, 0.5 * leaf._height
        rec['p'] = (wo2, ho2)
        if form == 0:
            xr, yr = x0 / dx, y0 / dy
            w, h = wo2 / abs(dx), ho2 / abs(dy)
            rec['i1'] = _clamp_ceil(-w - xr, 0, m)
            rec['i2'] = _clamp_ceil(w - xr, 0, m)
            rec['j1'] = _clamp_ceil(-h - yr, 0, n)
            rec['j2'] = _clamp_ceil(h - yr, 0, n)
        izero = jzero = 0
    elif isinstance(leaf, TopHat):
        # src/SBBox.cpp:268-370
        r0 = leaf._radius
        rec['type'] = _lib.LEAF_TOPHAT
        rec['norm'] = leaf._flux / (math.pi * (r0 * r0))
        rec['p'] = (r0 * r0, r0)
        if form == 0:
            rec['j1'] = _clamp_ceil(-r0 / abs(dy) - y0 / dy, 0, n)
            rec['j2'] = _clamp_ceil(r0 / abs(dy) - y0 / dy, 0, n)
        izero = jzero = 0
    else:
        raise NotImplementedError("drawReal of %s is not part of galsim_amd (Gaussian, "
                                  "Exponential, Moffat, Kolmogorov, Sersic, Airy, Spergel, Box, Pixel "
                                  "and TopHat are)"
                                  % type(leaf).__name__)
    if radial is not None:
        x0, dx, dxy, y0, dy, dyx = (x0 * radial, dx * radial, dxy * radial, y0 * radial,
                                    dy * radial, dyx * radial)
        if rec['type'] == _lib.LEAF_SERSIC and form == 1:
            # the pixel that is (0, 0) up to rounding takes r = 0 (src/SBSersic.cpp:185-212)
            det = dx * dy - dxy * dyx
            i0 = (-dy * x0 + dxy * y0) / det
            j0 = (dyx * x0 - dx * y0) / det
            ii, jj = int(math.floor(i0 + 0.5)), int(math.floor(j0 + 0.5))
            if abs(i0 - ii) < 1.e-12 and abs(j0 - jj) < 1.e-12 and 0 <= ii < m and 0 <= jj < n:
                rec['ic'], rec['jc'] = ii, jj
    rec['x0'], rec['dx'], rec['dxy'] = x0, dx, dxy
    rec['y0'], rec['dy'], rec['dyx'] = y0, dy, dyx
    rec['izero'], rec['jzero'] = (izero, jzero) if form == 0 else (0, 0)
    return rec


def components(prof, bounds, imscale=1., tables=None, jac=None, offset=(0., 0.),
               flux_scaling=1.):
    """gsmi_xcomp records (in drawing order) of ``prof`` drawn into an image with ``bounds``
    (its profile origin at (0, 0)) and pixel scale ``imscale``, under the Jacobian ``jac``, the
    offset and the flux ratio SBProfile::draw takes (none by default); the Kolmogorov radial
    tables the records index are appended to the list ``tables``."""
    if tables is None:
        tables = []
    from .convolve import Convolution
    from .sum import Sum
    from .transform import Transformation
    xmin, ymin = bounds.xmin, bounds.ymin
    m, n = bounds.xmax - bounds.xmin + 1, bounds.ymax - bounds.ymin + 1
    out = []

    def walk(obj, jac, offset, flux_scaling):
        if isinstance(obj, Transformation):
            dx, dy = offset
            if obj._has_offset:
                if jac is not None:
                    x1 = jac[0, 0] * obj._dx + jac[0, 1] * obj._dy
                    y1 = jac[1, 0] * obj._dx + jac[1, 1] * obj._dy
                else:
                    x1, y1 = obj._dx, obj._dy
                dx += x1
                dy += y1
            flux_scaling *= obj._flux_scaling
            jac = obj._jac if jac is None else jac if obj._jac is None else jac.dot(obj._jac)
            walk(obj._original, jac, (dx, dy), flux_scaling)
        elif isinstance(obj, Sum):
            for o in obj.obj_list:
                walk(o, jac, offset, flux_scaling)
        elif isinstance(obj, Convolution):
            if len(obj.obj_list) != 1:
                raise NotImplementedError("real-space convolution of %d profiles is not part of "
                                          "galsim_amd" % len(obj.obj_list))
            walk(obj.obj_list[0], jac, offset, flux_scaling)
        else:
            setup = _draw_setup(xmin, ymin, m, n, float(imscale), jac, float(offset[0]),
                                float(offset[1]), float(flux_scaling))
            out.append(_leaf_component(obj, m, n, setup, tables))

    walk(prof, None if jac is None else np.asarray(jac, dtype=float).reshape(2, 2),
         (float(offset[0]), float(offset[1])), float(flux_scaling))
    return np.array(out, dtype=XCOMP_DTYPE) if out else np.zeros(0, dtype=XCOMP_DTYPE)


def draw_real(prof, bounds, imscale=1., device_out=False, jac=None, offset=(0., 0.),
              flux_scaling=1.):
    """The FP64 image of ``prof`` sampled at the pixel centres of ``bounds`` (times the pixel
    area, as SBProfile::draw, with its Jacobian, offset and flux ratio if given), as a numpy
    array (or the device tensor with device_out)."""
    import torch
    if not torch.cuda.is_available():
        raise _lib.GalSimAMDError("galsim_amd drawReal runs on the GPU; no HIP device is "
                                  "available")
    tables = []
    recs = np.ascontiguousarray(components(prof, bounds, imscale, tables, jac, offset,
                                           flux_scaling))
    m, n = bounds.xmax - bounds.xmin + 1, bounds.ymax - bounds.ymin + 1
    out = torch.empty((n, m), dtype=torch.float64, device='cuda')
    dtab = (torch.from_numpy(np.concatenate(tables)).to('cuda') if tables else None)
    stream = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()
    if len(recs) == 0:
        out.zero_()
    for c0 in range(0, len(recs), XCOMP_MAX):
        chunk = np.ascontiguousarray(recs[c0:c0 + XCOMP_MAX])
        _lib.check(L.gsmi_draw_real_f64(chunk.ctypes.data, len(chunk),
                                        None if dtab is None else dtab.data_ptr(), m, n,
                                        int(c0 > 0), out.data_ptr(), stream),
                   "gsmi_draw_real_f64")
    return out if device_out else out.cpu().numpy()
