"""Oracle object layer and drawImage(method='fft') orchestration.

Test infrastructure only (see ``oracle/__init__.py``).

Restates, in numpy FP64:
* Transformation / _Transform composition and k-image drawing (galsim/transform.py:192-230,
  359-392, 490-517, 592-640), Shear.getMatrix (galsim/shear.py:365-386);
* Convolution and Sum (galsim/convolve.py:296-305, 420-426; galsim/sum.py:242-250, 341-347);
  Deconvolution, AutoConvolution, AutoCorrelation (galsim/convolve.py:464-935);
* SBProfile::drawK + FillQuadrant (src/SBProfile.cpp:383-506);
* ApplyKImagePhases with its renormalised phase recurrence (src/SBTransform.cpp:740-883);
* drawImage -> drawFFT -> drawFFT_makeKImage / drawFFT_finish (galsim/gsobject.py:1115-1253,
  1646-1844, 1900-2045) with PixelScale._profileToImage (galsim/wcs.py:1550-1558);
* wrapImage(hermx=True) as its closed form (tests/test_image.py:3388-3453 pins it) and
  irfft(shift_in, shift_out) as numpy.fft.irfft2 with the y-ifftshift and output fftshift
  (src/Image.cpp:814-911; galsim fft == numpy fft is pinned by tests/test_draw.py:1343-1420).
"""
import math

import numpy as np

from . import numerics as nm
from . import profiles as pr


# ----------------------------------------------------------------------------------------------
# Shear
# ----------------------------------------------------------------------------------------------
def shear_matrix(g1=None, g2=None, e1=None, e2=None):
    """Shear.getMatrix (galsim/shear.py:365-386); e->g by Shear._e2g (shear.py:324-330)."""
    if e1 is not None or e2 is not None:
        e1 = e1 or 0.
        e2 = e2 or 0.
        absesq = e1 * e1 + e2 * e2
        if absesq > 1.e-4:
            fac = 1. / (1. + np.sqrt(1. - absesq))
        else:
            fac = 0.5 + absesq * (0.125 + absesq * (0.0625 + absesq * 0.0390625))
        g = (e1 + 1j * e2) * fac
        g1, g2 = g.real, g.imag
    g1 = g1 or 0.
    g2 = g2 or 0.
    gabs = abs(g1 + 1j * g2)
    return np.array([[1. + g1, g2], [g2, 1. - g1]]) / np.sqrt(1. - gabs ** 2)


# ----------------------------------------------------------------------------------------------
# k-image drawing helpers
# ----------------------------------------------------------------------------------------------
class KImage:
    """A complex k-image over integer bounds (x fastest), scale dk."""

    def __init__(self, xmin, xmax, ymin, ymax, dk, dtype=np.complex128):
        self.xmin, self.xmax, self.ymin, self.ymax = xmin, xmax, ymin, ymax
        self.dk = dk
        self.array = np.zeros((ymax - ymin + 1, xmax - xmin + 1), dtype=dtype)

    def copy(self):
        k = KImage(self.xmin, self.xmax, self.ymin, self.ymax, self.dk, self.array.dtype)
        k.array = self.array.copy()
        return k


def draw_k(leaf, image, jac):
    """SBProfile::drawK (src/SBProfile.cpp:383-425) + the leaves' quadrant dispatch."""
    dk = image.dk
    n, m = image.array.shape
    x0 = image.xmin * dk
    izero = -image.xmin if image.xmin < 0 else 0
    y0 = image.ymin * dk
    jzero = -image.ymin if image.ymin < 0 else 0
    if jac is None:
        vals = _fill_1d_quadrant(leaf, x0, dk, izero, y0, dk, jzero, m, n)
    elif jac[0, 1] == 0. and jac[1, 0] == 0.:
        mA, mD = jac[0, 0], jac[1, 1]
        vals = _fill_1d_quadrant(leaf, x0 * mA, dk * mA, izero, y0 * mD, dk * mD, jzero, m, n)
    else:
        mA, mB, mC, mD = jac[0, 0], jac[0, 1], jac[1, 0], jac[1, 1]
        vals = leaf.fill_2d(mA * x0 + mC * y0, mA * dk, mC * dk,
                            mB * x0 + mD * y0, mD * dk, mB * dk, m, n)
    image.array[:, :] = vals


def _fill_1d_quadrant(leaf, x0, dx, m1, y0, dy, n1, m, n):
    """fillKImage(...izero, jzero) -> FillQuadrant (src/SBProfile.cpp:447-506)."""
    if m1 == 0 and n1 == 0:
        return leaf.fill_1d(x0, dx, m, y0, dy, n)
    m2 = m - m1 - 1
    n2 = n - n1 - 1
    q = leaf.fill_1d(x0 if m1 == 0 else 0., dx, max(m1, m2) + 1,
                     y0 if n1 == 0 else 0., dy, max(n1, n2) + 1)
    jj = np.abs(np.arange(n) - n1)
    ii = np.abs(np.arange(m) - m1)
    return q[jj][:, ii]


def _fillphase_1d(m, k, dk):
    """fillphase_1d (src/SBTransform.cpp:740-762): renormalised recurrence."""
    out = np.empty(m, dtype=np.complex128)
    kpol = complex(math.cos(-k), math.sin(-k))
    dkpol = complex(math.cos(-dk), math.sin(-dk))
    out[0] = kpol
    for i in range(1, m):
        kpol = kpol * dkpol
        kpol = kpol * (1.5 - 0.5 * (kpol.real * kpol.real + kpol.imag * kpol.imag))
        out[i] = kpol
    return out


def apply_kimage_phases(image, jac, cenx, ceny, flux_scaling):
    """ApplyKImagePhases (src/SBTransform.cpp:763-883)."""
    dk = image.dk
    n, m = image.array.shape
    x0 = image.xmin * dk
    y0 = image.ymin * dk
    if jac is None or (jac[0, 1] == 0. and jac[1, 0] == 0.):
        if jac is None:
            kx0, dkx, ky0, dky = x0, dk, y0, dk
        else:
            kx0, dkx, ky0, dky = x0 * jac[0, 0], dk * jac[0, 0], y0 * jac[1, 1], dk * jac[1, 1]
        kx0 *= cenx
        dkx *= cenx
        ky0 *= ceny
        dky *= ceny
        px = _fillphase_1d(m, kx0, dkx)
        py = _fillphase_1d(n, ky0, dky)
        image.array *= px[None, :] * (flux_scaling * py)[:, None]
    else:
        mA, mB, mC, mD = jac[0, 0], jac[0, 1], jac[1, 0], jac[1, 1]
        kx0 = (mA * x0 + mC * y0) * cenx
        dkx = mA * dk * cenx
        dkxy = mC * dk * cenx
        ky0 = (mB * x0 + mD * y0) * ceny
        dky = mD * dk * ceny
        dkyx = mB * dk * ceny
        k0 = kx0 + ky0
        dk0 = dkxy + dky
        dk1 = dkx + dkyx
        rows = pr._axis(k0, dk0, n)
        for j in range(n):
            image.array[j, :] *= flux_scaling * _fillphase_1d(m, rows[j], dk1)


# ----------------------------------------------------------------------------------------------
# Objects
# ----------------------------------------------------------------------------------------------
class Obj:
    """Minimal GSObject surface used by the tests (galsim/gsobject.py)."""

    # -- transformations (galsim/gsobject.py:840-1094) --
    def shear(self, g1=None, g2=None, e1=None, e2=None):
        return Transformation(self, jac=shear_matrix(g1, g2, e1, e2))

    def shift(self, dx, dy):
        return Transformation(self, offset=(dx, dy))

    def transform(self, dudx, dudy, dvdx, dvdy):
        return Transformation(self, jac=np.array([[dudx, dudy], [dvdx, dvdy]], dtype=float))

    def dilate(self, scale):
        return Transformation(self, jac=np.array([[scale, 0.], [0., scale]]),
                              flux_ratio=scale ** -2)

    def expand(self, scale):
        return Transformation(self, jac=np.array([[scale, 0.], [0., scale]]))

    def rotate(self, theta):
        c, s = math.cos(theta), math.sin(theta)
        return Transformation(self, jac=np.array([[c, -s], [s, c]]))

    def withScaledFlux(self, flux_ratio):
        return self if flux_ratio == 1 else Transformation(self, flux_ratio=flux_ratio)

    def __mul__(self, other):
        return self.withScaledFlux(other)

    __rmul__ = __mul__

    def __add__(self, other):
        return Sum([self, other])

    # -- sizes --
    @property
    def nyquist_scale(self):
        return math.pi / self.maxk

    def get_good_image_size(self, pixel_scale):
        """getGoodImageSize (galsim/gsobject.py:1900-1925)."""
        Nd = 2. * math.pi / (pixel_scale * self.stepk)
        N = int(math.ceil(Nd * (1. - 1.e-12)))
        return 2 * ((N + 1) // 2)

    # -- drawing --
    def drawImage(self, nx=None, ny=None, scale=None, method='fft', offset=(0., 0.),
                  use_true_center=True, area=1., exptime=1., gain=1., return_info=False):
        """drawImage for a new image (galsim/gsobject.py:1646-1844), FFT methods only."""
        if method not in ('fft', 'no_pixel', 'sb'):
            raise ValueError("oracle only restates the FFT methods")
        if scale is None or not scale > 0:
            scale = self.nyquist_scale
        ox, oy = float(offset[0]), float(offset[1])
        flux_scale = area * exptime
        if method == 'sb':
            flux_scale /= scale * scale
        if gain != 1:
            flux_scale /= gain
        # _adjust_offset (gsobject.py:1234-1253)
        if nx is not None:
            shape = (ny, nx)
        else:
            shape = (0, 0)
        if use_true_center:
            if shape[1] % 2 == 0:
                ox -= 0.5
            if shape[0] % 2 == 0:
                oy -= 0.5
        # PixelScale._profileToImage (wcs.py:1550-1558) with transform._Transform
        j = None if scale == 1. else np.array([[1. / scale, 0.], [0., 1. / scale]])
        prof = _Transform(self, j, (ox, oy), scale ** 2 * flux_scale)
        if method == 'fft':
            prof = Convolution([prof, Pixel(1.0, gsparams=self.gsparams)],
                               gsparams=self.gsparams)
        if nx is None:
            nx = ny = prof.get_good_image_size(1.0)
        # image bounds 1..nx, 1..ny, re-centred on (0,0) (gsobject.py:1775-1777)
        cx = 1 + nx // 2
        cy = 1 + ny // 2
        b = (1 - cx, nx - cx, 1 - cy, ny - cy)
        img, info = prof._draw_fft(b)
        img = img / 1.
        info['added_flux'] = info['added_photons'] / flux_scale
        if return_info:
            return img, info
        return img

    def _draw_fft(self, b):
        """drawFFT (gsobject.py:2014-2045) into bounds b (centred at 0)."""
        xmin, xmax, ymin, ymax = b
        N = self.get_good_image_size(1.0)
        image_N = max(max(abs(v) for v in b) * 2, ymax - ymin + 1, xmax - xmin + 1)
        N = max(N, image_N)
        N = nm.good_fft_size(N)
        N = max(N, self.gsparams.minimum_fft_size)
        dk = 2. * np.pi / (N * 1.0)
        maxk = self.maxk
        if N * dk / 2 > maxk:
            Nk = N
        else:
            Nk = int(np.ceil(maxk / dk)) * 2
        kimage = KImage(0, Nk // 2, -Nk // 2, Nk // 2, dk)
        self._drawKImage(kimage, None)
        kw = wrap_hermx(kimage.array, Nk, N)
        real = irfft_shifted(kw, N)
        # crop: real has bounds (-N/2..N/2-1)^2
        sub = real[ymin + N // 2: ymax + N // 2 + 1, xmin + N // 2: xmax + N // 2 + 1]
        info = dict(N=N, Nk=Nk, dk=dk, maxk=maxk, stepk=self.stepk,
                    added_photons=float(sub.sum(dtype=float)))
        return sub.copy(), info

    def drawKImage(self, nx, ny, scale):
        """drawKImage onto an nx x ny image with k-spacing ``scale`` centred at (0,0)
        (galsim/gsobject.py:2464-2583, the image-provided branch)."""
        cx = 1 + nx // 2
        cy = 1 + ny // 2
        kimage = KImage(1 - cx, nx - cx, 1 - cy, ny - cy, float(scale))
        self._drawKImage(kimage, None)
        return kimage.array


class Leaf(Obj):
    def __init__(self, prof):
        self.prof = prof
        self.gsparams = prof.gsparams

    @property
    def maxk(self):
        return self.prof.maxk

    @property
    def stepk(self):
        return self.prof.stepk

    @property
    def flux(self):
        return self.prof.flux

    def _drawKImage(self, image, jac):
        draw_k(self.prof, image, jac)


def Exponential(**kw):
    return Leaf(pr.Exponential(**kw))


def Gaussian(**kw):
    return Leaf(pr.Gaussian(**kw))


def Moffat(beta, **kw):
    return Leaf(pr.Moffat(beta, **kw))


def Kolmogorov(**kw):
    return Leaf(pr.Kolmogorov(**kw))


def Sersic(n, **kw):
    return Leaf(pr.Sersic(n, **kw))


def Spergel(nu, **kw):
    return Leaf(pr.Spergel(nu, **kw))


def Airy(lam_over_diam, **kw):
    return Leaf(pr.Airy(lam_over_diam, **kw))


def TopHat(radius, **kw):
    return Leaf(pr.TopHat(radius, **kw))


def DeltaFunction(**kw):
    return Leaf(pr.DeltaFunction(**kw))


def InclinedExponential(inclination, scale_radius, scale_height, **kw):
    return Leaf(pr.InclinedExponential(inclination, scale_radius, scale_height, **kw))


def InclinedSersic(n, inclination, scale_radius, scale_height, **kw):
    return Leaf(pr.InclinedSersic(n, inclination, scale_radius, scale_height, **kw))


def Box(width, height, flux=1., gsparams=None):
    return Leaf(pr.Box(width, height, flux, gsparams))


def Pixel(scale, flux=1., gsparams=None):
    return Leaf(pr.Pixel(scale, flux, gsparams))


class Transformation(Obj):
    """galsim/transform.py:192-230 (constructor composition) and 359-640."""

    def __init__(self, obj, jac=None, offset=(0., 0.), flux_ratio=1.):
        self._jac = None if jac is None else np.asarray(jac, dtype=float).reshape(2, 2)
        self._dx, self._dy = float(offset[0]), float(offset[1])
        self._flux_ratio = float(flux_ratio)
        self.gsparams = obj.gsparams
        if isinstance(obj, Transformation):
            if obj._has_offset:
                if jac is None:
                    dx1, dy1 = obj._dx, obj._dy
                else:
                    dx1, dy1 = _fwd_normal(self._jac, obj._dx, obj._dy)
                self._dx += dx1
                self._dy += dy1
            if jac is None:
                self._jac = obj._jac
            else:
                self._jac = self._jac if obj._jac is None else self._jac.dot(obj.jac)
            self._flux_ratio *= obj._flux_ratio
            self.original = obj.original
        else:
            self.original = obj
        self._has_offset = (self._dx != 0. or self._dy != 0.)

    @property
    def jac(self):
        return np.array([[1., 0.], [0., 1.]]) if self._jac is None else self._jac

    @property
    def _det(self):
        j = self._jac
        if j is None:
            return 1
        if j[0, 1] == 0. and j[1, 0] == 0.:
            if j[0, 0] == 1. and j[1, 1] == 1.:
                return 1.
            return j[0, 0] * j[1, 1]
        return j[0, 0] * j[1, 1] - j[0, 1] * j[1, 0]

    @property
    def _flux_scaling(self):
        return abs(self._det) * self._flux_ratio

    @property
    def flux(self):
        return self._flux_scaling * self.original.flux

    def _major_minor(self):
        j = self._jac
        if j is None:
            return 1., 1.
        h1 = math.hypot(j[0, 0] + j[1, 1], j[0, 1] - j[1, 0])
        h2 = math.hypot(j[0, 0] - j[1, 1], j[0, 1] + j[1, 0])
        return 0.5 * abs(h1 + h2), 0.5 * abs(h1 - h2)

    @property
    def maxk(self):
        return self.original.maxk / self._major_minor()[1]

    @property
    def stepk(self):
        stepk = self.original.stepk / self._major_minor()[0]
        if self._dx != 0. or self._dy != 0.:
            dr = math.hypot(self._dx, self._dy)
            stepk = math.pi / (math.pi / stepk + dr)
        return stepk

    def _drawKImage(self, image, jac):
        jac1 = self._jac if jac is None else jac if self._jac is None else jac.dot(self._jac)
        self.original._drawKImage(image, jac1)
        if self._has_offset:
            apply_kimage_phases(image, jac, self._dx, self._dy, self._flux_scaling)
        elif abs(self._flux_scaling - 1.) > self.gsparams.kvalue_accuracy:
            image.array *= self._flux_scaling


def _fwd_normal(jac, x, y):
    """Transformation._fwd_normal (galsim/transform.py:436-444)."""
    temp = jac[0, 1] * y
    y = y * jac[1, 1]
    y += jac[1, 0] * x
    x = x * jac[0, 0]
    x += temp
    return x, y


def _Transform(obj, jac=None, offset=(0., 0.), flux_ratio=1.):
    """transform._Transform (galsim/transform.py:604-640)."""
    ret = Transformation.__new__(Transformation)
    ret.gsparams = obj.gsparams
    ret._jac = jac
    ret._dx, ret._dy = offset
    if isinstance(obj, Transformation):
        if obj._has_offset:
            if jac is None:
                dx1, dy1 = obj._dx, obj._dy
            else:
                dx1, dy1 = _fwd_normal(ret._jac, obj._dx, obj._dy)
            ret._dx += dx1
            ret._dy += dy1
        if jac is None:
            ret._jac = obj._jac
        else:
            ret._jac = ret._jac if obj._jac is None else ret._jac.dot(obj.jac)
        ret._flux_ratio = flux_ratio * obj._flux_ratio
        ret.original = obj.original
    else:
        ret._flux_ratio = flux_ratio
        ret.original = obj
    ret._has_offset = (ret._dx != 0. or ret._dy != 0.)
    return ret


class Convolution(Obj):
    """galsim/convolve.py:124-230, 296-305, 420-426 (Fourier convolution only)."""

    def __init__(self, obj_list, gsparams=None):
        self.obj_list = list(obj_list)
        self.gsparams = gsparams or pr.GSParams.combine([o.gsparams for o in self.obj_list])

    @property
    def maxk(self):
        return min(o.maxk for o in self.obj_list)

    @property
    def stepk(self):
        return sum(o.stepk ** (-2) for o in self.obj_list) ** (-0.5)

    @property
    def flux(self):
        return float(np.prod([o.flux for o in self.obj_list]))

    def _drawKImage(self, image, jac):
        self.obj_list[0]._drawKImage(image, jac)
        if len(self.obj_list) > 1:
            im1 = image.copy()
            for obj in self.obj_list[1:]:
                obj._drawKImage(im1, jac)
                image.array *= im1.array


def Convolve(*objs, **kw):
    if len(objs) == 1 and isinstance(objs[0], (list, tuple)):
        objs = objs[0]
    return Convolution(objs, **kw)


class Deconvolution(Obj):
    """galsim/convolve.py:464-612: the original's k-image inverted where its modulus exceeds
    flux * kvalue_accuracy (else the inverse of that threshold), zeroed beyond its maxk."""

    def __init__(self, obj):
        self.orig = obj
        self.gsparams = obj.gsparams
        self.min_acc = obj.flux * self.gsparams.kvalue_accuracy
        self.inv_min_acc = 1. / self.min_acc

    @property
    def maxk(self):
        return self.orig.maxk

    @property
    def stepk(self):
        return self.orig.stepk

    @property
    def flux(self):
        return 1. / self.orig.flux

    def _drawKImage(self, image, jac):
        self.orig._drawKImage(image, jac)
        a = image.array
        do_inverse = np.abs(a) > self.min_acc
        a[do_inverse] = 1. / a[do_inverse]
        a[~do_inverse] = self.inv_min_acc
        kx, ky = np.meshgrid(np.arange(a.shape[1], dtype=float) + image.xmin,
                             np.arange(a.shape[0], dtype=float) + image.ymin)
        if jac is not None:
            kx, ky = (jac[0, 0] * kx + jac[1, 0] * ky), (jac[0, 1] * kx + jac[1, 1] * ky)
        ksq = (kx ** 2 + ky ** 2) * image.dk ** 2
        a[ksq > self.maxk ** 2] = 0.


def Deconvolve(obj):
    return Deconvolution(obj)


def AutoConvolve(obj):
    """galsim/convolve.py:646-771: a Convolution of [obj, obj]."""
    return Convolution([obj, obj], gsparams=obj.gsparams)


def AutoCorrelate(obj):
    """galsim/convolve.py:805-935: a Convolution of [obj, obj.transform(-1, 0, 0, -1)]."""
    return Convolution([obj, obj.transform(-1, 0, 0, -1)], gsparams=obj.gsparams)


class Sum(Obj):
    """galsim/sum.py:242-250, 341-347."""

    def __init__(self, obj_list, gsparams=None):
        self.obj_list = list(obj_list)
        self.gsparams = gsparams or pr.GSParams.combine([o.gsparams for o in self.obj_list])

    @property
    def maxk(self):
        return float(np.max([o.maxk for o in self.obj_list]))

    @property
    def stepk(self):
        return float(np.min([o.stepk for o in self.obj_list]))

    @property
    def flux(self):
        return float(np.sum([o.flux for o in self.obj_list]))

    def _drawKImage(self, image, jac):
        self.obj_list[0]._drawKImage(image, jac)
        if len(self.obj_list) > 1:
            im1 = image.copy()
            for obj in self.obj_list[1:]:
                obj._drawKImage(im1, jac)
                image.array += im1.array


def Add(*objs):
    if len(objs) == 1 and isinstance(objs[0], (list, tuple)):
        objs = objs[0]
    return Sum(objs)


# ----------------------------------------------------------------------------------------------
# Wrap and inverse FFT
# ----------------------------------------------------------------------------------------------
def wrap_hermx(karr, Nk, N):
    """wrapImage(hermx=True, hermy=False) of the (0..Nk/2, -Nk/2..Nk/2) k-image onto
    (0..N/2, -N/2..N/2-1) (src/Image.cpp:577-726), written as its closed form: every point of the
    implicit full plane (x<0 from Hermitian symmetry) is added to the target cell congruent to
    it modulo N (tests/test_image.py:3388-3453).  Returns an (N, N/2+1) array, row 0 = ky=-N/2.
    """
    half = Nk // 2
    kx = np.arange(0, half + 1)
    ky = np.arange(-half, half + 1)
    KX, KY = np.meshgrid(kx, ky)
    out = np.zeros((N, N // 2 + 1), dtype=np.complex128)
    # stored points (kx >= 0)
    c = KX % N
    r = (KY + N // 2) % N
    sel = c <= N // 2
    np.add.at(out, (r[sel], c[sel]), karr[sel])
    # mirrored points (-kx, -ky) for kx > 0
    c2 = (-KX) % N
    r2 = (-KY + N // 2) % N
    sel2 = (KX > 0) & (c2 <= N // 2)
    np.add.at(out, (r2[sel2], c2[sel2]), np.conj(karr[sel2]))
    return out


def irfft_shifted(kw, N):
    """irfft(kwrap, real, shift_in=True, shift_out=True) (src/Image.cpp:814-911): the real
    image on (-N/2..N/2-1)^2 (the two padding columns are not returned)."""
    shifted = np.fft.ifftshift(kw, axes=0)
    real = np.fft.irfft2(shifted, s=(N, N))
    return np.fft.fftshift(real)
