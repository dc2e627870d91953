"""CPU restatement of the reference's real-space drawing (drawImage method='no_pixel' / 'sb' of
analytic profiles).  Test infrastructure only: the checker of galsim_amd.real, never called by it.

draw_real(obj, xmin, xmax, ymin, ymax, imscale) follows GSObject.drawReal -> _drawReal of
Transformation / Sum / one-object Convolution (galsim/gsobject.py:1846-1898,
galsim/transform.py:571-583, galsim/sum.py:304-310, galsim/convolve.py:395-397) into
SBProfile::draw (src/SBProfile.cpp:304-381) and the leaves' fillXImage loops, with their
accumulated coordinates (x += dx), the quadrant fill (src/SBProfile.cpp:447-488) and the
component-by-component sum, in float64.

Pinned by the reference's own saved images (tests/SBProfile_comparison_images, 5 decimals, in
tests/test_real_cpu.py) and by closed forms (flux sums, peak values, symmetric profiles drawn
through shear / rotation matrices).
"""
import math

import numpy as np
import scipy.special as sps

from . import numerics as nm
from .draw import Convolution, Leaf, Sum, Transformation
from .profiles import (Airy, Box, Exponential, Gaussian, Kolmogorov, Moffat, Sersic, Spergel,
                       TopHat, fast_pow, kolmogorov_xtable)


def _acc(x0, dx, m):
    """x0, x0 + dx, ... as the reference's loops accumulate them."""
    out = np.empty(m)
    x = x0
    for i in range(m):
        out[i] = x
        x += dx
    return out


def _grid(x0, dx, dxy, y0, dy, dyx, m, n):
    """Coordinates of the general fill: rows start at accumulated (x0 + j dxy, y0 + j dy), then
    x += dx, y += dyx along the row."""
    X = np.empty((n, m))
    Y = np.empty((n, m))
    X[:, 0] = _acc(x0, dxy, n)
    Y[:, 0] = _acc(y0, dy, n)
    for i in range(1, m):
        X[:, i] = X[:, i - 1] + dx
        Y[:, i] = Y[:, i - 1] + dyx
    return X, Y


class _X:
    """x-space constants of one leaf (each leaf's constructor in the reference)."""

    def __init__(self, prof):
        self.prof = prof
        if isinstance(prof, Gaussian):
            # src/SBGaussian.cpp:55-76, 95-190
            self.inv = 1. / prof.sigma
            self.norm = prof.flux * (self.inv * self.inv) / (2. * math.pi)
            self.kind = 'gauss'
        elif isinstance(prof, Exponential):
            # src/SBExponential.cpp:70-90, 381-440
            self.inv = 1. / prof.r0
            self.norm = (prof.flux / (2. * math.pi)) * (self.inv * self.inv)
            self.kind = 'exp'
        elif isinstance(prof, Moffat):
            # src/SBMoffat.cpp:170-272, 340-415
            b = prof.beta
            self.inv = 1. / prof.rD
            if prof.trunc > 0.:
                maxRrD = prof.trunc * self.inv
                ff = 1. - math.pow(1. + maxRrD * maxRrD, 1. - b)
            else:
                ff = 1.
                maxRrD = math.sqrt(math.pow(prof.gsparams.xvalue_accuracy, 1. / (1. - b)) - 1.)
            self.maxRrD_sq = maxRrD * maxRrD
            self.norm = prof.flux * (b - 1.) / (math.pi * ff * (prof.rD * prof.rD))
            xacc = prof.gsparams.xvalue_accuracy
            self.pw = lambda x: fast_pow(x, -b)
            for s, f in ((1., lambda x: 1. / x), (1.5, lambda x: 1. / (x * np.sqrt(x))),
                         (2., lambda x: 1. / (x * x)), (2.5, lambda x: 1. / (x * x * np.sqrt(x))),
                         (3., lambda x: 1. / (x * x * x)),
                         (3.5, lambda x: 1. / (x * x * x * np.sqrt(x))),
                         (4., lambda x: 1. / ((x * x) * (x * x)))):
                if abs(b - s) < xacc:
                    self.pw = f
                    break
            self.kind = 'moffat'
        elif isinstance(prof, Sersic):
            # src/SBSersic.cpp:70-222, SersicInfo::getXNorm / xValue (354-360)
            info = prof.info
            self.inv = 1. / prof.r0
            xn = 1. / (2. * math.pi * info.n * info.gamma2n * info.flux_fraction)
            self.norm = (prof.flux * xn) * (self.inv * self.inv)
            self.inv2n = info.inv2n
            self.trunc_sq = info.trunc * info.trunc if info.trunc > 0. else 0.
            self.kind = 'sersic'
        elif isinstance(prof, Airy):
            # src/SBAiry.cpp:66-80, AiryInfoObs / AiryInfoNoObs radial functions 88-108, 410-429
            self.inv = 1. / prof.lod
            self.norm = prof.flux * (self.inv * self.inv)
            self.obs = prof.obs
            self.thresh = math.sqrt(8. * prof.gsparams.xvalue_accuracy)
            self.kind = 'airy'
        elif isinstance(prof, Kolmogorov):
            # src/SBKolmogorov.cpp:83-115, KolmogorovInfo::xValue 114-115
            self.inv = prof.k0
            self.norm = prof.flux * (prof.k0 * prof.k0)
            args, vals = kolmogorov_xtable(prof.gsparams)
            self.table = nm.SplineTable(args, vals)
            self.rmax = args[-1]
            self.kind = 'kolm'
        elif isinstance(prof, Spergel):
            # src/SBSpergel.cpp:80-95, 525-535, SpergelInfo::getXNorm / xValue 644-651
            nu = prof.nu
            g1 = math.gamma(nu + 1.)
            self.inv = 1. / prof.r0
            self.norm = (prof.flux * (math.pow(2., -nu) / g1 / (2.0 * math.pi))) / (
                prof.r0 * prof.r0)
            self.nu = nu
            self.xnorm0 = g1 / (2. * nu) * math.pow(2., nu) if nu > 0. else math.inf
            self.kind = 'spergel'
        elif isinstance(prof, Box):
            # src/SBBox.cpp:55-148
            self.norm = prof.flux / (prof.width * prof.height)
            self.wo2, self.ho2 = 0.5 * prof.width, 0.5 * prof.height
            self.kind = 'box'
        elif isinstance(prof, TopHat):
            # src/SBBox.cpp:268-370
            self.r0 = prof.r0
            self.r0sq = prof.r0 * prof.r0
            self.norm = prof.flux / (math.pi * self.r0sq)
            self.kind = 'tophat'
        else:
            raise NotImplementedError(type(prof).__name__)

    def radial(self, rsq):
        with np.errstate(divide='ignore', over='ignore', invalid='ignore'):
            if self.kind == 'exp':
                return self.norm * nm.expd(-np.sqrt(rsq))
            if self.kind == 'moffat':
                return np.where(rsq <= self.maxRrD_sq, self.norm * self.pw(1. + rsq), 0.)
            if self.kind == 'sersic':
                v = self.norm * nm.expd(-fast_pow(rsq, self.inv2n))
                if self.trunc_sq > 0.:
                    v = np.where(rsq > self.trunc_sq, 0., v)
                return v
            if self.kind == 'gauss':
                return self.norm * nm.expd(-0.5 * rsq)
            if self.kind == 'kolm':
                r = np.sqrt(rsq)
                flat = r.ravel()
                v = np.array([self.table.lookup(float(q)) if q < self.rmax else 0.
                              for q in flat]).reshape(r.shape)
                return self.norm * v
            if self.kind == 'spergel':
                r = np.sqrt(rsq)
                v = sps.kv(self.nu, r) * fast_pow(r, self.nu)
                return self.norm * np.where(r == 0., self.xnorm0, v)
            if self.kind == 'airy':
                e = self.obs
                nu = np.sqrt(rsq) * math.pi
                xv = np.where(nu < self.thresh, 0.5 * (1. - e * e),
                              (sps.j1(nu) - e * sps.j1(e * nu)) / np.where(nu > 0., nu, 1.))
                if e == 0.:
                    xv = np.where(nu < self.thresh, 0.5, sps.j1(nu) / np.where(nu > 0., nu, 1.))
                    return self.norm * ((xv * xv) * math.pi)
                return self.norm * ((xv * xv) * (math.pi / (1. - e * e)))
        raise AssertionError(self.kind)

    def fill_plain(self, x0, dx, m, y0, dy, n):
        """The 4-argument fill without the quadrant split (izero = jzero = 0)."""
        x0, dx, y0, dy = x0 * self.inv, dx * self.inv, y0 * self.inv, dy * self.inv
        xs = _acc(x0, dx, m)
        ys = _acc(y0, dy, n)
        if self.kind == 'gauss':
            gx = nm.expd(-0.5 * xs * xs)
            gy = nm.expd(-0.5 * ys * ys)
            return (self.norm * gx)[None, :] * gy[:, None]
        return self.radial(xs[None, :] * xs[None, :] + (ys * ys)[:, None])

    def fill4(self, x0, dx, izero, y0, dy, jzero, m, n):
        if self.kind == 'box':
            xr, yr = x0 / dx, y0 / dy
            w, h = self.wo2 / abs(dx), self.ho2 / abs(dy)
            i1 = max(0, math.ceil(-w - xr))
            i2 = min(m, math.ceil(w - xr))
            j1 = max(0, math.ceil(-h - yr))
            j2 = min(n, math.ceil(h - yr))
            out = np.zeros((n, m))
            if i2 > i1 and j2 > j1:
                out[j1:j2, i1:i2] = self.norm
            return out
        if self.kind == 'tophat':
            out = np.zeros((n, m))
            adx, ady = abs(dx), abs(dy)
            j1 = max(0, math.ceil(-self.r0 / ady - y0 / dy))
            j2 = min(n, math.ceil(self.r0 / ady - y0 / dy))
            y = y0 + j1 * dy
            for j in range(j1, j2):
                d = self.r0sq - y * y
                if d >= 0.:
                    xmax = math.sqrt(d)
                    i1 = max(0, math.ceil(-xmax / adx - x0 / dx))
                    i2 = min(m, math.ceil(xmax / adx - x0 / dx))
                    if i2 > i1:
                        out[j, i1:i2] = self.norm
                y += dy
            return out
        if izero == 0 and jzero == 0:
            return self.fill_plain(x0, dx, m, y0, dy, n)
        # FillQuadrant: one quadrant from the origin outwards, mirrored
        m1, n1 = izero, jzero
        m2, n2 = m - m1 - 1, n - n1 - 1
        q = self.fill_plain(x0 if m1 == 0 else 0., dx, max(m1, m2) + 1,
                            y0 if n1 == 0 else 0., dy, max(n1, n2) + 1)
        ii = np.abs(np.arange(m) - m1)
        jj = np.abs(np.arange(n) - n1)
        return q[np.ix_(jj, ii)]

    def fill6(self, x0, dx, dxy, y0, dy, dyx, m, n):
        if self.kind in ('box', 'tophat'):
            X, Y = _grid(x0, dx, dxy, y0, dy, dyx, m, n)
            if self.kind == 'box':
                outside = (np.abs(X) > self.wo2) | (np.abs(Y) > self.ho2)
                inside = (np.abs(X) < self.wo2) & (np.abs(Y) < self.ho2)
            else:
                rsq = X * X + Y * Y
                outside = rsq > self.r0sq
                inside = rsq < self.r0sq
            out = np.zeros((n, m))
            for j in range(n):
                i = 0
                while i < m and outside[j, i]:
                    i += 1
                while i < m and inside[j, i]:
                    out[j, i] = self.norm
                    i += 1
            return out
        s = self.inv
        x0, dx, dxy, y0, dy, dyx = x0 * s, dx * s, dxy * s, y0 * s, dy * s, dyx * s
        X, Y = _grid(x0, dx, dxy, y0, dy, dyx, m, n)
        out = self.radial(X * X + Y * Y)
        if self.kind == 'sersic':
            det = dx * dy - dxy * dyx
            i0 = (-dy * x0 + dxy * y0) / det
            j0 = (dyx * x0 - dx * y0) / det
            ii, jj = int(math.floor(i0 + 0.5)), int(math.floor(j0 + 0.5))
            if abs(i0 - ii) < 1.e-12 and abs(j0 - jj) < 1.e-12 and 0 <= ii < m and 0 <= jj < n:
                out[jj, ii] = self.norm
        return out


def draw_leaf(prof, xmin, ymin, m, n, imscale, jac, xoff, yoff, flux_ratio):
    """SBProfile::draw (src/SBProfile.cpp:304-381) of one leaf into an (n, m) array."""
    xc = _X(prof)
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
        img = xc.fill4(x0, imscale, izero, y0, imscale, jzero, m, n)
        if imscale != 1.:
            flux_ratio *= imscale * imscale
    elif jac[0, 1] == 0. and jac[1, 0] == 0.:
        xscale, yscale = 1. / jac[0, 0], 1. / jac[1, 1]
        dx, dy = imscale * xscale, imscale * yscale
        img = xc.fill4(x0 * xscale, dx, izero, y0 * yscale, dy, jzero, m, n)
        flux_ratio *= abs(dx * dy)
    else:
        mA, mB, mC, mD = jac[0, 0], jac[0, 1], jac[1, 0], jac[1, 1]
        invdet = 1. / (mA * mD - mB * mC)
        nx0 = (mD * x0 - mB * y0) * invdet
        ny0 = (-mC * x0 + mA * y0) * invdet
        invdet *= imscale
        img = xc.fill6(nx0, mD * invdet, -mB * invdet, ny0, mA * invdet, -mC * invdet, m, n)
        flux_ratio *= abs(invdet * imscale)
    if flux_ratio != 1.:
        img = img * flux_ratio
    return img


def draw_real(obj, xmin, xmax, ymin, ymax, imscale=1.):
    """The (ny, nx) float64 image of drawReal over bounds [xmin, xmax] x [ymin, ymax]."""
    m, n = xmax - xmin + 1, ymax - ymin + 1
    parts = []

    def walk(o, jac, offset, flux_scaling):
        if isinstance(o, Transformation):
            dx, dy = offset
            if o._dx != 0. or o._dy != 0.:
                if jac is not None:
                    dx += jac[0, 0] * o._dx + jac[0, 1] * o._dy
                    dy += jac[1, 0] * o._dx + jac[1, 1] * o._dy
                else:
                    dx += o._dx
                    dy += o._dy
            flux_scaling *= o._flux_scaling
            jac = o._jac if jac is None else jac if o._jac is None else jac.dot(o._jac)
            walk(o.original, jac, (dx, dy), flux_scaling)
        elif isinstance(o, Sum):
            for p in o.obj_list:
                walk(p, jac, offset, flux_scaling)
        elif isinstance(o, Convolution):
            assert len(o.obj_list) == 1
            walk(o.obj_list[0], jac, offset, flux_scaling)
        elif isinstance(o, Leaf):
            parts.append(draw_leaf(o.prof, xmin, ymin, m, n, imscale, jac, offset[0], offset[1],
                                   flux_scaling))
        else:
            raise NotImplementedError(type(o).__name__)

    walk(obj, None, (0., 0.), 1.)
    img = parts[0]
    for p in parts[1:]:
        img = img + p
    return img


def draw_image_sb(obj, xmin, xmax, ymin, ymax, scale, method='sb', offset=(0., 0.)):
    """drawImage(image with these bounds centred at (0, 0), scale, method, use_true_center=False)
    of an analytic profile: the PixelScale transformation (galsim/wcs.py:1550-1558) then
    drawReal with unit pixels."""
    from .draw import _Transform
    flux_scale = 1. / (scale * scale) if method == 'sb' else 1.
    j = None if scale == 1. else np.array([[1. / scale, 0.], [0., 1. / scale]])
    prof = _Transform(obj, j, (float(offset[0]), float(offset[1])), scale ** 2 * flux_scale)
    return draw_real(prof, xmin, xmax, ymin, ymax, 1.)
