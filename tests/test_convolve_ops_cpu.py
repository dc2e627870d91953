"""Pins for the oracle's Deconvolution, AutoConvolution and AutoCorrelation (oracle/draw.py):
the identities of the reference's own tests (tests/test_convolve.py:477-700) and the explicit
k-space behaviour of the deconvolution (inverse, threshold, maxk cut)."""
import math

import numpy as np
import pytest

import galsim_amd as g
from oracle import draw as od
from oracle import profiles as pr


def _psf():
    return od.Moffat(3.8, fwhm=1.3, flux=5., gsparams=pr.GSParams(maxk_threshold=3.e-4))


def _asym():
    o1 = od.Gaussian(sigma=3., flux=4.).shift(-0.2, -0.4)
    o2 = od.Gaussian(sigma=6., flux=1.3).shift(0.3, 0.3)
    return od.Add(o1, o2)


def test_deconvolve_cancels_a_factor():
    # tests/test_convolve.py:482-500: Convolve(psf, psf, Deconvolve(psf)) draws as psf
    psf = _psf()
    a = od.Convolve([psf, psf, od.Deconvolve(psf)]).drawImage(80, 80, scale=0.4,
                                                               method='no_pixel')
    b = psf.drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(a, b, 4)
    assert od.Deconvolve(psf).flux == pytest.approx(1. / psf.flux)
    # tests/test_convolve.py:508-522: an asymmetric sum of shifted Gaussians
    obj = _asym()
    a = od.Convolve([od.Deconvolve(obj), obj, obj]).drawImage(80, 80, scale=0.4,
                                                               method='no_pixel')
    b = obj.drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(a, b, 4)
    # tests/test_convolve.py:530-541: transformed
    jac = (0.3, -0.8, -0.7, 0.4)
    tobj = obj.transform(*jac)
    tinv = od.Deconvolve(obj).transform(*jac)
    tinv = tinv * (1. / (tobj.flux * tinv.flux))
    a = od.Convolve([tinv, tobj, obj]).drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(a, b, 4)


def test_deconvolution_kimage_values():
    # the k-image is 1/F inside the original's maxk where |F| > flux * kvalue_accuracy, the
    # inverse threshold where |F| is below it, and exactly 0 beyond maxk
    psf = od.Moffat(2.5, fwhm=1.3, flux=2., gsparams=pr.GSParams(maxk_threshold=1.e-7))
    inv = od.Deconvolve(psf)
    dk = 0.9 * psf.maxk / 20.
    F = od.KImage(0, 40, -40, 40, dk)
    psf._drawKImage(F, None)
    K = od.KImage(0, 40, -40, 40, dk)
    inv._drawKImage(K, None)
    x, y = np.meshgrid(np.arange(0, 41), np.arange(-40, 41))
    k = np.hypot(x, y) * dk
    inside = k <= psf.maxk * (1 - 1e-12)
    beyond = k > psf.maxk * (1 + 1e-12)
    big = np.abs(F.array) > 2. * 1e-5
    small = np.abs(F.array) < 2. * 1e-5 * 0.999
    assert beyond.any() and (inside & small).any() and (inside & big).any()
    np.testing.assert_array_equal(K.array[beyond], 0.)
    np.testing.assert_allclose(K.array[inside & big] * F.array[inside & big], 1., rtol=1e-14)
    np.testing.assert_array_equal(K.array[inside & small], 1. / (2. * 1e-5))


def test_autoconvolve_and_autocorrelate_identities():
    psf = od.Moffat(3.8, fwhm=1.3, flux=5.)
    a = od.AutoConvolve(psf).drawImage(80, 80, scale=0.4, method='no_pixel')
    b = od.Convolve([psf, psf]).drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_equal(a, b)
    c = od.AutoCorrelate(psf).drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(c, b, 4)         # symmetric: the same thing
    # tests/test_convolve.py:675-700: AutoCorrelate(add) = add convolved with its 180-degree
    # rotation
    add = _asym()
    c = od.AutoCorrelate(add).drawImage(80, 80, scale=0.8, method='no_pixel')
    d = od.Convolve([add, add.rotate(math.pi)]).drawImage(80, 80, scale=0.8, method='no_pixel')
    np.testing.assert_array_almost_equal(c, d, 4)
    # and not the same as the auto-convolution for this asymmetric profile
    e = od.AutoConvolve(add).drawImage(80, 80, scale=0.8, method='no_pixel')
    assert np.max(np.abs(c - e)) > 1e-3 * np.max(np.abs(d))


def test_product_argument_checks():
    # tests/test_convolve.py:555-567, 650-660
    psf = g.Moffat(beta=3.8, fwhm=1.3, flux=5)
    for f in (g.Deconvolve, g.Deconvolution, g.AutoConvolve, g.AutoConvolution,
              g.AutoCorrelate, g.AutoCorrelation):
        with pytest.raises(TypeError):
            f()
        with pytest.raises(TypeError):
            f([psf])
        with pytest.raises(TypeError):
            f(psf, psf)
    with pytest.raises(TypeError):
        g.Deconvolve(psf, real_space=False)
    inv = g.Deconvolve(psf)
    assert inv.flux == pytest.approx(1. / psf.flux)
    assert inv.maxk == psf.maxk and inv.stepk == psf.stepk
    ac = g.AutoConvolve(psf)
    assert ac.flux == pytest.approx(psf.flux ** 2)
    assert ac.stepk == pytest.approx(psf.stepk / math.sqrt(2.), rel=1e-14)
    assert g.AutoCorrelate(psf).maxk == psf.maxk
