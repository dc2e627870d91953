"""InterpolatedImage on the GPU against the oracle: drawImage through the separable and the
cross-term k fill, every interpolant as x and k interpolant, FP64 (1e-10 of the peak pixel) and
FP32 (1e-5), kValue against the reference's saved values, drawKImage, and batches mixing
InterpolatedImages with analytic profiles."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import galsim_amd as g
from galsim_amd import batch
from oracle import draw as od
from oracle import interpimage as ii
from tests.test_interpimage_cpu import KX, KY, GOLD, ref_image

F64_TOL = 1e-10
F32_TOL = 1e-5


@pytest.fixture(scope='module')
def ref():
    return ref_image()


def _peak_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _pair(arr, scale, **kw):
    prod = g.InterpolatedImage(g.Image(arr), scale=scale, **kw)
    orc = ii.InterpolatedImage(arr, scale, **kw)
    return prod, orc


def _check(prod, orc, nx, ny, scale, precision='f64', **kw):
    print("case: %s nx=%d ny=%d scale=%r %s %r" % (prod, nx, ny, scale, precision, kw),
          flush=True)
    im = prod.drawImage(nx=nx, ny=ny, scale=scale, dtype=np.float64, precision=precision, **kw)
    okw = dict(kw)
    if okw.get('method') == 'auto':
        okw['method'] = 'fft'            # the pixel convolution, by FFT, as 'auto' takes it here
    want, info = orc.drawImage(nx, ny, scale=scale, return_info=True, **okw)
    err = _peak_rel(im.array, want)
    print("  N %d Nk %d peak-rel %.2e" % (info['N'], info['Nk'], err))
    assert err < (F64_TOL if precision == 'f64' else F32_TOL), err
    return im, want


@pytest.mark.parametrize("xi,ki", [(None, None), ('lanczos3', 'cubic'), ('linear', 'lanczos5F'),
                                   ('cubic', 'quintic'), ('nearest', 'linear'),
                                   ('lanczos7T', 'lanczos4')])
def test_draw_matches_oracle_interpolants(ref, xi, ki):
    arr = ref[32:96, 28:100]                            # 72 x 64, off-centre content
    prod, orc = _pair(arr, 0.4, x_interpolant=xi, k_interpolant=ki, normalization='sb')
    assert prod.stepk == orc.stepk and prod.flux == orc.flux
    assert abs(prod.maxk - orc.maxk) <= 1e-15 * orc.maxk
    _check(prod, orc, 64, 64, 0.4, offset=(0.3, -0.2))


def test_draw_scales_offsets_flux(ref):
    arr = ref[40:88, 40:88]
    prod, orc = _pair(arr, 0.4, flux=3.0, offset=(0.5, 0.25), use_true_center=False)
    _check(prod, orc, 48, 40, 0.3, method='auto')
    _check(prod, orc, 33, 35, 0.55, method='fft')
    prod, orc = _pair(arr, 1.0, normalization='flux')
    _check(prod, orc, 40, 40, 1.0)


def test_sheared_and_convolved_takes_the_cross_term_fill(ref):
    """A shear on the InterpolatedImage makes the transforms hand it a non-diagonal k grid (the
    fill with cross terms, cut at |k| > maxk1); convolved with a Moffat PSF."""
    arr = ref[36:92, 36:92]
    prod, orc = _pair(arr, 0.4, x_interpolant='lanczos3')
    p = g.Convolve([prod.shear(g1=0.2, g2=-0.1).shift(0.1, 0.05),
                    g.Moffat(beta=3., fwhm=0.9)])
    o = od.Convolve(orc.shear(g1=0.2, g2=-0.1).shift(0.1, 0.05), od.Moffat(3., fwhm=0.9))
    _check(p, o, 64, 64, 0.35)
    p = prod.rotate(0.7 * g.radians).dilate(1.3)
    o = orc.rotate(0.7).dilate(1.3)
    _check(p, o, 56, 56, 0.4)


def test_fp32_draw(ref):
    arr = ref[32:96, 32:96]
    prod, orc = _pair(arr, 0.4, normalization='sb')
    _check(g.Convolve([prod, g.Kolmogorov(fwhm=0.8)]),
           od.Convolve(orc, od.Kolmogorov(fwhm=0.8)), 64, 64, 0.4, precision='f32')
    _check(prod.shear(g1=-0.15, g2=0.2), orc.shear(g1=-0.15, g2=0.2), 64, 64, 0.4,
           precision='f32')


@pytest.mark.parametrize("name,xi", [('Cubic', 'cubic'), ('Quintic', 'quintic'),
                                     ('Lanczos5', 'lanczos5F'), ('Lanczos7', 'lanczos7F')])
def test_kvalue_matches_oracle_and_saved_values(ref, name, xi):
    prod, orc = _pair(ref, 0.4, x_interpolant=xi, normalization='sb')
    r = np.loadtxt('%s/absfK%s_test.txt' % (GOLD, name))
    got = np.array([prod.kValue(kx, ky) for kx, ky in zip(KX, KY)])
    want = np.array([orc.kValue(kx, ky) for kx, ky in zip(KX, KY)])
    print(name, got, want)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * abs(orc.kValue(0., 0.)))
    np.testing.assert_allclose(r / np.abs(got), 1., atol=1e-6)


def test_drawkimage_matches_oracle(ref):
    arr = ref[40:88, 40:88]
    prod, orc = _pair(arr, 0.4, x_interpolant='cubic', offset=(0.2, 0.1))
    kim = prod.drawKImage(nx=33, ny=29, scale=0.3)
    want = orc.drawKImage(33, 29, 0.3)
    err = float(np.max(np.abs(kim.array - want)) / np.max(np.abs(want)))
    print("drawKImage peak-rel", err)
    assert err < 1e-12


@pytest.mark.parametrize("precision", ['f64', 'f32'])
def test_batch_mixes_interpolated_images_and_profiles(ref, precision):
    """A batch of InterpolatedImages (two images, shared k-images), sheared copies and analytic
    profiles against per-object drawImage (FP64) and the oracle."""
    a1 = ref[32:96, 32:96]
    a2 = ref[44:84, 40:90]
    i1 = g.InterpolatedImage(g.Image(a1), scale=0.4)
    i2 = g.InterpolatedImage(g.Image(a2), scale=0.4, x_interpolant='lanczos3')
    o1 = ii.InterpolatedImage(a1, 0.4)
    o2 = ii.InterpolatedImage(a2, 0.4, x_interpolant='lanczos3')
    psf, opsf = g.Kolmogorov(fwhm=0.8), od.Kolmogorov(fwhm=0.8)
    profs = [g.Convolve([i1, psf]), g.Convolve([i2.shear(g1=0.1, g2=0.2), psf]),
             g.Convolve([g.Exponential(half_light_radius=1.), psf]), g.Convolve([i1, psf])]
    oracle = [od.Convolve(o1, opsf), od.Convolve(o2.shear(g1=0.1, g2=0.2), opsf),
              od.Convolve(od.Exponential(half_light_radius=1.), opsf), od.Convolve(o1, opsf)]
    nx = np.array([48, 40, 32, 48])
    offs = np.array([[0.1, -0.2], [0., 0.3], [0., 0.], [0.25, 0.25]])
    plan = batch.BatchPlan(profs, nx, nx, 0.4, offsets=offs, precision=precision)
    plan.run()
    tol = F64_TOL if precision == 'f64' else F32_TOL
    for i, (o, n) in enumerate(zip(oracle, nx)):
        want = o.drawImage(int(n), int(n), scale=0.4, offset=tuple(offs[i]))
        got = plan.stamp(i).cpu().numpy()
        err = _peak_rel(got, want)
        print("stamp", i, precision, "peak-rel %.2e" % err)
        assert err < tol, (i, err)
    assert not np.any(plan.status())


def test_unsupported_forms_raise(ref):
    arr = ref[40:88, 40:88]
    prod = g.InterpolatedImage(g.Image(arr), scale=0.4)
    with pytest.raises(NotImplementedError):
        prod.drawImage(nx=32, ny=32, scale=0.4, method='real_space')
    with pytest.raises(NotImplementedError):
        g.InterpolatedImage(g.Image(arr), scale=0.4, x_interpolant='sinc')
    with pytest.raises(NotImplementedError):
        g.InterpolatedImage(g.Image(arr), scale=0.4, noise_pad_size=10, noise_pad=0.1)
    with pytest.raises(ValueError):
        g.InterpolatedImage(g.Image(np.zeros((8, 8))), scale=0.4)
