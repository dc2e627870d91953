"""The _galsim-shaped boundary (galsim_amd._galsim), driven the way GalSim's own Python layer
drives its _galsim module: profile objects whose ``_sbp`` is built through the SB constructors,
the ``_drawKImage`` recursion of galsim/gsobject.py, galsim/transform.py:592-600,
galsim/convolve.py:420-426 and galsim/sum.py:341-347 (leaves call ``drawK``, shifts call
``ApplyKImagePhases``), then drawFFT_finish's ``wrapImage`` + ``irfft`` + crop
(galsim/gsobject.py:1975-2012).  Compared with the oracle's drawImage at 1e-10 of the peak."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import galsim_amd as g
from galsim_amd import _galsim as G
from galsim_amd import render
from galsim_amd.batch import _image_profile
from oracle import draw as od

F64_TOL = 1e-10
GSP = G.GSParams(128, 8192, 5e-3, 5, 1e-3, 1e-5, 1e-5, 1, 1e-4, 1e-6, 1e-6, 1e-8, 1e-5)
KVALUE_ACCURACY = 1e-5


def _addr(a):
    return 0 if a is None else a.__array_interface__['data'][0]


class KImage:
    """A complex128 image on (xmin..xmax, ymin..ymax) with its _galsim view (galsim.Image)."""

    def __init__(self, xmin, xmax, ymin, ymax, scale, array=None):
        self.b = (xmin, xmax, ymin, ymax)
        self.scale = scale
        self.array = (np.zeros((ymax - ymin + 1, xmax - xmin + 1), dtype=np.complex128)
                      if array is None else array)
        self._image = G.ImageViewCD(self.array.ctypes.data, 1, self.array.strides[0] // 16,
                                    G.BoundsI(*self.b))

    def copy(self):
        return KImage(*self.b, self.scale, self.array.copy())


# galsim's Python objects, reduced to what drawFFT touches: _sbp built through _galsim, and
# _drawKImage as galsim defines it
class Leaf:
    def __init__(self, sbp):
        self._sbp = sbp

    def _drawKImage(self, image, jac=None):                 # GSObject._drawKImage
        self._sbp.drawK(image._image, image.scale, _addr(jac))


class Transformation:
    def __init__(self, original, jac, offset, flux_ratio):
        self._original = original
        self._jac = None if jac is None else np.ascontiguousarray(jac, dtype=float)
        self._dx, self._dy = offset
        self._flux_ratio = flux_ratio
        j = self._jac
        det = 1. if j is None else abs(j[0, 0] * j[1, 1] - j[0, 1] * j[1, 0])
        self._flux_scaling = det * flux_ratio                # galsim/transform.py:359-392
        self._sbp = G.SBTransform(original._sbp, _addr(self._jac), self._dx, self._dy,
                                  flux_ratio, GSP)

    def _drawKImage(self, image, jac=None):                 # galsim/transform.py:592-600
        jac1 = self._jac if jac is None else jac if self._jac is None else jac.dot(self._jac)
        self._original._drawKImage(image, jac1)
        if self._dx != 0. or self._dy != 0.:
            G.ApplyKImagePhases(image._image, image.scale, _addr(jac), self._dx, self._dy,
                                self._flux_scaling)
        elif abs(self._flux_scaling - 1.) > KVALUE_ACCURACY:
            image.array *= self._flux_scaling


class Convolution:
    def __init__(self, objs):
        self.obj_list = objs
        self._sbp = G.SBConvolve([o._sbp for o in objs], False, GSP)

    def _drawKImage(self, image, jac=None):                 # galsim/convolve.py:420-426
        self.obj_list[0]._drawKImage(image, jac)
        if len(self.obj_list) > 1:
            tmp = image.copy()
            for obj in self.obj_list[1:]:
                obj._drawKImage(tmp, jac)
                image.array *= tmp.array


class Sum:
    def __init__(self, objs):
        self.obj_list = objs
        self._sbp = G.SBAdd([o._sbp for o in objs], GSP)

    def _drawKImage(self, image, jac=None):                 # galsim/sum.py:341-347
        self.obj_list[0]._drawKImage(image, jac)
        if len(self.obj_list) > 1:
            tmp = image.copy()
            for obj in self.obj_list[1:]:
                obj._drawKImage(tmp, jac)
                image.array += tmp.array


def to_galsim_style(obj):
    """The galsim-style object of a galsim_amd profile tree (the same parameters)."""
    if isinstance(obj, g.Transformation):
        return Transformation(to_galsim_style(obj._original), obj._jac, (obj._dx, obj._dy),
                              obj._flux_ratio)
    if isinstance(obj, g.Convolution):
        return Convolution([to_galsim_style(o) for o in obj.obj_list])
    if isinstance(obj, g.Sum):
        return Sum([to_galsim_style(o) for o in obj.obj_list])
    if isinstance(obj, g.Exponential):
        return Leaf(G.SBExponential(obj.scale_radius, obj.flux, GSP))
    if isinstance(obj, g.Moffat):
        return Leaf(G.SBMoffat(obj.beta, obj.scale_radius, obj.trunc, obj.flux, GSP))
    if isinstance(obj, g.Sersic):
        return Leaf(G.SBSersic(obj.n, obj.scale_radius, obj.flux, obj.trunc, GSP))
    if isinstance(obj, g.Kolmogorov):
        return Leaf(G.SBKolmogorov(obj.lam_over_r0, obj.flux, GSP))
    if isinstance(obj, g.Box):
        return Leaf(G.SBBox(obj.width, obj.height, obj.flux, GSP))
    if isinstance(obj, g.InterpolatedImage):
        return interpolated_image_style(obj)
    raise TypeError(type(obj))


def _galsim_interp(i):
    if isinstance(i, g.Lanczos):
        return G.Lanczos(i.n, i.conserve_dc, GSP)
    return getattr(G, type(i).__name__)(GSP)


def interpolated_image_style(obj):
    """InterpolatedImage._sbp as galsim builds it (galsim/interpolatedimage.py:398-427): an
    SBInterpolatedImage of the padded image with stepk / maxk in pixel units, shifted by
    -offset, then the pixel scale's transform with the flux ratio (galsim/wcs.py:1538-1548)."""
    xim = np.ascontiguousarray(obj._xim)
    N = xim.shape[0]
    view = G.ImageViewD(xim.ctypes.data, 1, N, G.BoundsI(-(N // 2), N // 2 - 1, -(N // 2),
                                                        N // 2 - 1))
    b = G.BoundsI(obj._bx0, obj._bx0 + obj._nx - 1, obj._by0, obj._by0 + obj._ny - 1)
    s = obj._scale
    sbii = G.SBInterpolatedImage(view, b, b, _galsim_interp(obj.x_interpolant),
                                 _galsim_interp(obj.k_interpolant), obj.stepk * s,
                                 obj.maxk * s, GSP)
    leaf = Leaf(sbii)
    leaf._keep = xim
    ratio = obj.flux / obj._image_flux if obj.flux != obj._image_flux else 1.
    shifted = Transformation(leaf, None, (-obj._offset[0], -obj._offset[1]), 1.)
    return Transformation(shifted, np.array([[s, 0.], [0., s]]), (0., 0.),
                          (1. / s) ** 2 * ratio)


def draw_fft_via_galsim_boundary(prof, nx, ny, scale, offset):
    """drawImage(method='fft') as galsim performs it through _galsim: the image-coordinate
    profile (PixelScale transform, offset, pixel), drawFFT_makeKImage, _drawKImage,
    drawFFT_finish."""
    p, b, _ = _image_profile(prof, nx, ny, scale, offset, True, 'fft')
    N, Nk, dk = render.plan_sizes(p, b)
    obj = to_galsim_style(p)
    kimage = KImage(0, Nk // 2, -(Nk // 2), Nk // 2, dk)
    obj._drawKImage(kimage)
    k_direct = KImage(0, Nk // 2, -(Nk // 2), Nk // 2, dk)
    obj._sbp.drawK(k_direct._image, dk, 0)                  # the C++-side tree fill
    filled = kimage.copy()                                  # wrapImage folds kimage in place
    # drawFFT_finish
    G.wrapImage(kimage._image, G.BoundsI(0, N // 2, -(N // 2), N // 2 - 1), True, False)
    r0 = Nk // 2 - N // 2
    kwrap = kimage.array[r0:r0 + N, :N // 2 + 1]
    kview = G.ImageViewCD(kwrap.ctypes.data, 1, kimage.array.strides[0] // 16,
                          G.BoundsI(0, N // 2, -(N // 2), N // 2 - 1))
    real = np.empty((N, N + 2))
    G.irfft(kview, G.ImageViewD(real.ctypes.data, 1, N + 2,
                                G.BoundsI(-(N // 2), N // 2 + 1, -(N // 2), N // 2 - 1)),
            True, True)
    img = real[b.ymin + N // 2:b.ymax + N // 2 + 1, b.xmin + N // 2:b.xmax + N // 2 + 1]
    return img, filled, k_direct, (N, Nk)


def _peak_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _case(m, kind):
    if kind == 'exp_moffat':
        gal = m.Exponential(half_light_radius=1.5, flux=1e4).shear(g1=0.1, g2=-0.05)
        return m.Convolve([gal, m.Moffat(beta=3, fwhm=0.7)])
    if kind == 'shifted_sum_sheared':
        # a shift inside a sheared transform: ApplyKImagePhases's general-Jacobian branch
        gal = (m.Exponential(half_light_radius=1.0, flux=3e3).shift(0.3, -0.2) +
               m.Sersic(2.5, half_light_radius=0.6, flux=2e3)).shear(g1=0.2, g2=0.15)
        return m.Convolve([gal, m.Kolmogorov(fwhm=0.8)])
    if kind == 'moffat_trunc_rotated':
        gal = m.Moffat(beta=2.5, half_light_radius=0.9, trunc=4.0, flux=5e3)
        gal = gal.rotate(0.4 * (g.radians if m is g else 1.)).shift(-0.15, 0.1)
        return m.Convolve([gal, m.Kolmogorov(fwhm=0.7)])
    raise ValueError(kind)


@pytest.mark.parametrize("kind,offset", [('exp_moffat', (0.13, -0.21)),
                                         ('shifted_sum_sheared', (-0.3, 0.25)),
                                         ('moffat_trunc_rotated', (0.0, 0.0))])
def test_drawfft_through_galsim_boundary_matches_oracle(kind, offset):
    print("case", kind, offset, flush=True)
    img, kimage, k_direct, (N, Nk) = draw_fft_via_galsim_boundary(_case(g, kind), 64, 64, 0.2,
                                                                   offset)
    ref = _case(od, kind).drawImage(64, 64, scale=0.2, offset=offset)
    err = _peak_rel(img, ref)
    print("N", N, "Nk", Nk, "peak-rel vs oracle", err)
    assert err < F64_TOL
    # the SB tree's own drawK (the C++ SBConvolve/SBTransform/SBAdd fill) gives the same k-image
    # as galsim's Python recursion; compare before the wrap touched kimage: redraw it
    k_py = KImage(0, Nk // 2, -(Nk // 2), Nk // 2, kimage.scale)
    p, _, _ = _image_profile(_case(g, kind), 64, 64, 0.2, offset, True, 'fft')
    to_galsim_style(p)._drawKImage(k_py)
    assert np.max(np.abs(k_direct.array - k_py.array)) < 1e-12 * np.max(np.abs(k_py.array))


def test_galsim_boundary_scalars_match_library():
    e = G.SBExponential(1.3, 100., GSP)
    ref = g.Exponential(scale_radius=1.3, flux=100.)
    assert e.maxK() == ref.maxk and e.stepK() == ref.stepk and e.getFlux() == 100.
    assert G.goodFFTSize(1174) == 1536
    k = e.kValue(G.PositionD(0.7, -0.2))
    ksq = (0.7 ** 2 + 0.2 ** 2) * 1.3 ** 2
    assert abs(k - 100. / (1 + ksq) ** 1.5) < 1e-12 * 100.


def test_apply_kimage_phases_closed_form():
    # ApplyKImagePhases multiplies by fs exp(-i k'.c), k' = J^T k, k = (x, y) dk
    rng = np.random.default_rng(3)
    im = KImage(-3, 5, -4, 4, 0.37)
    im.array[...] = rng.normal(size=im.array.shape) + 1j * rng.normal(size=im.array.shape)
    before = im.array.copy()
    jac = np.array([[1.1, 0.3], [-0.2, 0.9]])
    G.ApplyKImagePhases(im._image, im.scale, _addr(jac), 0.4, -0.7, 2.5)
    y, x = np.mgrid[-4:5, -3:6] * 0.37
    kxp = jac[0, 0] * x + jac[1, 0] * y
    kyp = jac[0, 1] * x + jac[1, 1] * y
    want = before * 2.5 * np.exp(-1j * (kxp * 0.4 + kyp * -0.7))
    np.testing.assert_allclose(im.array, want, rtol=0, atol=1e-13)


def test_interpolated_image_through_galsim_boundary():
    """An InterpolatedImage built as galsim builds it on _galsim (SBInterpolatedImage, its
    calculateMaxK, the shift and pixel-scale transforms), drawn by drawFFT through the boundary,
    against the oracle's InterpolatedImage; CalculateSizeContainingFlux and calculateMaxK against
    the set-up of galsim_amd.InterpolatedImage."""
    from oracle import interpimage as ii
    from tests.test_interpimage_cpu import ref_image
    arr = ref_image()[36:92, 40:90]
    for xi in ('quintic', 'lanczos3'):
        prod = g.InterpolatedImage(g.Image(arr), scale=0.3, x_interpolant=xi, offset=(0.2, -0.1))
        orc = ii.InterpolatedImage(arr, 0.3, x_interpolant=xi, offset=(0.2, -0.1))
        # the set-up helpers of the boundary give the product's numbers
        xim = np.ascontiguousarray(prod._xim)
        N = xim.shape[0]
        b = G.BoundsI(prod._bx0, prod._bx0 + prod._nx - 1, prod._by0, prod._by0 + prod._ny - 1)
        im_view = G.ImageViewD(xim[b.ymin + N // 2:, b.xmin + N // 2:].ctypes.data, 1, N, b)
        R = G.CalculateSizeContainingFlux(im_view, (1. - 5e-3) * prod._image_flux)
        assert math.pi / (math.hypot(R, prod.x_interpolant.xrange) * 0.3) == prod.stepk
        view = G.ImageViewD(xim.ctypes.data, 1, N, G.BoundsI(-(N // 2), N // 2 - 1,
                                                            -(N // 2), N // 2 - 1))
        sb = G.SBInterpolatedImage(view, b, b, _galsim_interp(prod.x_interpolant),
                                   _galsim_interp(prod.k_interpolant), prod.stepk * 0.3, 0., GSP)
        sb.calculateMaxK(0.)
        assert sb.maxK() / 0.3 == prod.maxk
        # drawFFT through the boundary against the oracle
        print("case InterpolatedImage", xi, flush=True)
        img, kimage, k_direct, (N2, Nk) = draw_fft_via_galsim_boundary(prod, 48, 48, 0.3,
                                                                        (0.1, 0.05))
        want = orc.drawImage(48, 48, scale=0.3, offset=(0.1, 0.05))
        err = _peak_rel(img, want)
        kerr = float(np.max(np.abs(kimage.array - k_direct.array)) / np.max(np.abs(
            k_direct.array)))
        print("N", N2, "Nk", Nk, "peak-rel vs oracle", err, "k-image recursion vs tree", kerr)
        assert err < F64_TOL and kerr < 1e-12


def test_real_space_draw_and_xvalue_through_galsim_boundary():
    """SBProfile.draw (src/SBProfile.cpp:304-381) on a view, as galsim's _drawReal calls it
    on a leaf (Jacobian, offset, flux ratio), against the oracle's real-space restatement; and
    xValue against the closed form of a sheared, shifted Gaussian."""
    import oracle.xdraw as ox
    sb = G.SBExponential(0.7, 2.5, GSP)
    arr = np.zeros((21, 24))
    view = G.ImageViewD(arr.ctypes.data, 1, 24, G.BoundsI(-12, 11, -10, 10))
    jac = np.ascontiguousarray([[1.1, 0.2], [-0.1, 0.9]])
    sb.draw(view, 0.2, _addr(jac), 0.13, -0.07, 1.7)
    want = ox.draw_real(od.Transformation(od.Exponential(scale_radius=0.7, flux=2.5), jac,
                                          (0.13, -0.07), 1.7 / abs(np.linalg.det(jac))),
                        -12, 11, -10, 10, 0.2)
    err = _peak_rel(arr, want)
    print("draw peak-rel vs oracle", err)
    assert err < 1e-12
    gs = G.SBTransform(G.SBGaussian(0.8, 3.0, GSP), _addr(np.ascontiguousarray(
        [[1.2, 0.1], [0.1, 0.9]])), 0.2, -0.1, 1.0, GSP)
    x, y = 0.35, -0.4
    J = np.array([[1.2, 0.1], [0.1, 0.9]])
    u = np.linalg.solve(J, np.array([x - 0.2, y + 0.1]))
    # SBTransform's amplitude scaling multiplies the surface brightness (its flux by |det J|)
    val = 3.0 / (2. * np.pi * 0.8 ** 2) * np.exp(-0.5 * (u @ u) / 0.8 ** 2)
    got = gs.xValue(G.PositionD(x, y))
    print("xValue", got, val)
    assert abs(got - val) < 1e-13 * val
