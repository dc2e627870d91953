"""GPU parity of Deconvolution, AutoConvolution and AutoCorrelation against the oracle (FP64:
1e-10 of the peak pixel; FP32: 1e-5), with the identities of the reference's tests
(tests/test_convolve.py:477-700)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import galsim_amd as g
from oracle import draw as od
from oracle import profiles as pr

F64_TOL = 1e-10
F32_TOL = 1e-5


def _peak_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _psf(m, thr=3.e-4):
    if m is g:
        return g.Moffat(beta=3.8, fwhm=1.3, flux=5).withGSParams(maxk_threshold=thr)
    return od.Moffat(3.8, fwhm=1.3, flux=5., gsparams=pr.GSParams(maxk_threshold=thr))


def _asym(m):
    o1 = m.Gaussian(sigma=3., flux=4.).shift(-0.2, -0.4)
    o2 = m.Gaussian(sigma=6., flux=1.3).shift(0.3, 0.3)
    return m.Add(o1, o2)


def _transformed(m):
    obj = _asym(m)
    jac = (0.3, -0.8, -0.7, 0.4)
    tobj = obj.transform(*jac)
    tinv = m.Deconvolve(obj).transform(*jac)
    tinv = tinv * (1. / (tobj.flux * tinv.flux))
    return m.Convolve([tinv, tobj, obj])


CASES = {
    'psf2_over_psf': lambda m: m.Convolve([_psf(m), _psf(m), m.Deconvolve(_psf(m))]),
    'asym_deconv': lambda m: m.Convolve([m.Deconvolve(_asym(m)), _asym(m), _asym(m)]),
    'transformed_deconv': _transformed,
    'autoconvolve_moffat': lambda m: m.AutoConvolve(_psf(m, 1.e-3)),
    'autocorrelate_asym': lambda m: m.AutoCorrelate(_asym(m)),
    'autoconvolve_asym_sheared': lambda m: m.AutoConvolve(_asym(m)).shear(g1=0.2, g2=-0.1),
    'deconv_shifted_exp': lambda m: m.Convolve(
        [m.Exponential(half_light_radius=1.5, flux=3.).shear(g1=0.1, g2=0.2),
         m.Deconvolve(m.Gaussian(sigma=0.4).shift(0.1, -0.05))]),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("method", ['no_pixel', 'fft'])
def test_drawimage_matches_oracle(name, method):
    print("case", name, method, flush=True)
    make = CASES[name]
    im = make(g).drawImage(nx=80, ny=80, scale=0.4, method=method, dtype=np.float64)
    ref, info = make(od).drawImage(80, 80, scale=0.4, method=method, return_info=True)
    print("N", info['N'], "Nk", info['Nk'], "peak-rel", _peak_rel(im.array, ref))
    assert _peak_rel(im.array, ref) < F64_TOL
    im32 = make(g).drawImage(nx=80, ny=80, scale=0.4, method=method, dtype=np.float32,
                             precision='f32')
    print("f32 peak-rel", _peak_rel(im32.array.astype(float), ref))
    assert _peak_rel(im32.array.astype(float), ref) < F32_TOL


def test_reference_identities():
    # tests/test_convolve.py:482-541 and 569-700 at the reference's 4 decimals
    # (the single analytic profiles draw in real space in the reference: their images here come
    # from the oracle's k-space drawing)
    psf = _psf(g)
    a = g.Convolve(psf, psf, g.Deconvolve(psf)).drawImage(nx=80, ny=80, scale=0.4,
                                                          method='no_pixel')
    b = _psf(od).drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(a.array, b, 4)
    obj = _asym(g)
    a = g.Convolve([g.Deconvolve(obj), obj, obj]).drawImage(nx=80, ny=80, scale=0.4,
                                                            method='no_pixel')
    b = _asym(od).drawImage(80, 80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(a.array, b, 4)
    a = _transformed(g).drawImage(nx=80, ny=80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(a.array, b, 4)
    m = g.Moffat(beta=3.8, fwhm=1.3, flux=5)
    a = g.AutoConvolve(m).drawImage(nx=80, ny=80, scale=0.4, method='no_pixel')
    b = g.Convolve([m, m]).drawImage(nx=80, ny=80, scale=0.4, method='no_pixel')
    np.testing.assert_array_equal(a.array, b.array)
    c = g.AutoCorrelate(m).drawImage(nx=80, ny=80, scale=0.4, method='no_pixel')
    np.testing.assert_array_almost_equal(c.array, b.array, 4)
    add = _asym(g)
    c = g.AutoCorrelate(add).drawImage(nx=80, ny=80, scale=0.8, method='no_pixel')
    d = g.Convolve([add, add.rotate(math.pi)]).drawImage(nx=80, ny=80, scale=0.8,
                                                        method='no_pixel')
    np.testing.assert_array_almost_equal(c.array, d.array, 4)


def test_deconvolution_kimage_and_kvalue():
    psf_g = g.Moffat(beta=2.5, fwhm=1.3, flux=2.).withGSParams(maxk_threshold=1.e-7)
    psf_o = od.Moffat(2.5, fwhm=1.3, flux=2., gsparams=pr.GSParams(maxk_threshold=1.e-7))
    dk = 0.9 * psf_g.maxk / 20.
    kim = g.ImageCD(81, 81, scale=dk)
    g.Deconvolve(psf_g).drawKImage(kim)
    ref = od.Deconvolve(psf_o).drawKImage(81, 81, dk)
    assert np.max(np.abs(kim.array - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert (ref == 0).any() and (np.abs(ref - 1. / (2. * 1e-5)) < 1e-6).any()
    # kValue: the inverse with the threshold, no maxk cut
    for kx, ky in ((0.3, -0.2), (1.1, 0.4)):
        v = g.Deconvolve(psf_g).kValue(kx, ky)
        f = psf_g.kValue(kx, ky)
        assert v == pytest.approx(1. / f, rel=1e-14)
