"""GPU parity of the HIP drawing path against the oracle (FP64: 1e-10 of the peak pixel;
FP32: 1e-5 of the peak pixel)."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import galsim_amd as g
from galsim_amd import render
from oracle import draw as od
from tests.fitsio import read_fits

F64_TOL = 1e-10
F32_TOL = 1e-5


def _both(make):
    return make(g), make(od)


def _peak_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _draw_pair(make, nx, ny, scale=0.2, precision='f64', **kw):
    # method 'fft' (the oracle's restatement): under 'auto' a lone hard-edged profile would take
    # the reference's real-space convolution, which this library raises on
    kw.setdefault('method', 'fft')
    prod, orc = _both(make)
    im = prod.drawImage(nx=nx, ny=ny, scale=scale, dtype=np.float64, precision=precision, **kw)
    ref, info = orc.drawImage(nx, ny, scale=scale, return_info=True, **kw)
    return im, ref, info


def _incl(m, inc, r0, h0, flux=1.):
    if m is g:
        return g.InclinedExponential(inc * g.radians, scale_radius=r0, scale_height=h0, flux=flux)
    return od.InclinedExponential(inc, r0, h0, flux=flux)


def _incls(m, n, inc, r0, h0, flux=1., trunc=0.):
    if m is g:
        return g.InclinedSersic(n, inc * g.radians, scale_radius=r0, scale_height=h0, flux=flux,
                                trunc=trunc)
    return od.InclinedSersic(n, inc, r0, h0, flux=flux, trunc=trunc)


def _expmoffat(m, hlr=1.5, g1=0.1, g2=-0.05, beta=3, fwhm=0.7):
    gal = m.Exponential(half_light_radius=hlr, flux=1e4).shear(g1=g1, g2=g2)
    return m.Convolve([gal, m.Moffat(beta=beta, fwhm=fwhm)])


def test_exponential_moffat_stamp0():
    # Exponential(hlr=1.5", flux=1e4) sheared (0.1, -0.05), offset (0.13, -0.21) px, with
    # Moffat(beta=3, fwhm=0.7") on a 64x64 stamp at 0.2"/px
    im, ref, info = _draw_pair(lambda m: _expmoffat(m), 64, 64, offset=(0.13, -0.21))
    print("N", info['N'], "Nk", info['Nk'], "peak-rel", _peak_rel(im.array, ref))
    assert _peak_rel(im.array, ref) < F64_TOL
    assert abs(im.added_flux - info['added_flux']) < 1e-10 * info['added_flux']


def test_exponential_moffat_seeded_sweep():
    rng = np.random.default_rng(1234)
    worst = 0.
    for _ in range(12):
        hlr = rng.uniform(0.8, 2.5)
        gmag = rng.uniform(0, 0.5)
        phi = rng.uniform(0, np.pi)
        g1, g2 = gmag * np.cos(2 * phi), gmag * np.sin(2 * phi)
        off = tuple(rng.uniform(-0.5, 0.5, 2))
        im, ref, info = _draw_pair(lambda m: _expmoffat(m, hlr, g1, g2), 64, 64, offset=off)
        worst = max(worst, _peak_rel(im.array, ref))
    print("worst peak-rel", worst)
    assert worst < F64_TOL


@pytest.mark.parametrize("name,make,shape", [
    ("gaussian_sheared", lambda m: m.Gaussian(sigma=0.7, flux=3.).shear(e1=0.3, e2=0.1),
     (40, 40)),
    ("kolmogorov_folded", lambda m: m.Kolmogorov(fwhm=0.6, flux=2.), (32, 32)),
    ("moffat_15", lambda m: m.Moffat(1.5, scale_radius=0.4), (48, 48)),
    ("moffat_25_shift", lambda m: m.Moffat(2.5, half_light_radius=0.5).shift(0.1, -0.2),
     (48, 48)),
    ("moffat_generic", lambda m: m.Moffat(2.7, fwhm=0.9).shear(g1=-0.2, g2=0.15), (48, 48)),
    ("moffat_2", lambda m: m.Moffat(2., fwhm=0.6, flux=4.), (48, 48)),
    ("moffat_37_sheared", lambda m: m.Moffat(3.7, fwhm=0.8).shear(g1=0.1, g2=0.2), (48, 48)),
    ("moffat_44_conv", lambda m: m.Convolve([m.Exponential(half_light_radius=0.7, flux=9.),
                                             m.Moffat(4.4, fwhm=0.9)]), (56, 56)),
    ("moffat_3_sheared", lambda m: m.Moffat(3., fwhm=0.7).shear(g1=0.25, g2=-0.1), (40, 44)),
    ("moffat_4_conv", lambda m: m.Convolve([m.Exponential(half_light_radius=0.5, flux=30.),
                                            m.Moffat(4., fwhm=0.8)]), (48, 48)),
    ("sersic_15", lambda m: m.Sersic(1.5, half_light_radius=0.6, flux=5.), (48, 48)),
    ("sersic_4_conv", lambda m: m.Convolve([m.Sersic(4., half_light_radius=0.8),
                                            m.Kolmogorov(fwhm=0.8)]), (64, 64)),
    ("bulge_disk", lambda m: m.Convolve([
        (m.Exponential(half_light_radius=1.0, flux=600.) +
         m.Sersic(2.5, half_light_radius=0.5, flux=400.)).shear(g1=0.3, g2=-0.2),
        m.Kolmogorov(fwhm=0.75)]), (64, 64)),
    ("moffat_trunc_2", lambda m: m.Moffat(2., fwhm=0.9, trunc=1.8, flux=3.), (48, 48)),
    ("moffat_trunc_105_sheared", lambda m: m.Moffat(1.05, scale_radius=0.3, trunc=1.5)
     .shear(g1=0.15, g2=-0.1), (48, 48)),
    ("moffat_trunc_conv", lambda m: m.Convolve([m.Exponential(half_light_radius=0.6, flux=20.),
                                                m.Moffat(3.2, half_light_radius=0.5, trunc=2.)]),
     (56, 56)),
    ("sersic_trunc_conv", lambda m: m.Convolve([m.Sersic(2.5, half_light_radius=0.6,
                                                         trunc=2.0, flux=7.),
                                                m.Kolmogorov(fwhm=0.7)]), (56, 56)),
    ("sersic_trunc_sheared", lambda m: m.Sersic(1.5, scale_radius=0.3, trunc=1.5)
     .shear(g1=-0.2, g2=0.25), (48, 48)),
    ("odd_nonsquare", lambda m: _expmoffat(m, 1.1, 0.2, 0.3), (37, 52)),
    ("exp_small_clipped", lambda m: m.Exponential(scale_radius=0.02).shear(g1=0.4, g2=0.1),
     (24, 24)),
    ("spergel_02", lambda m: m.Spergel(0.2, half_light_radius=0.8, flux=17.9), (48, 48)),
    ("spergel_m06_conv", lambda m: m.Convolve([m.Spergel(-0.6, half_light_radius=0.6, flux=3.)
                                               .shear(g1=0.2, g2=-0.1),
                                               m.Kolmogorov(fwhm=0.7)]), (56, 56)),
    ("spergel_3_shift", lambda m: m.Spergel(3., scale_radius=0.15).shift(0.1, 0.05),
     (40, 40)),
    ("spergel_small_clipped", lambda m: m.Spergel(0.85, scale_radius=0.03)
     .shear(g1=0.3, g2=0.2), (24, 24)),
    ("airy", lambda m: m.Airy(lam_over_diam=0.6, flux=2.), (40, 40)),
    ("airy_obs_sheared", lambda m: m.Airy(lam_over_diam=0.8, obscuration=0.32)
     .shear(g1=0.2, g2=0.1), (48, 48)),
    ("airy_obs_conv", lambda m: m.Convolve([m.Exponential(half_light_radius=0.7, flux=9.),
                                            m.Airy(lam_over_diam=0.5, obscuration=0.1)]),
     (48, 48)),
    ("inclined_13", lambda m: _incl(m, 1.3, 0.6, 0.1, 10.), (48, 48)),
    ("inclined_edge_on_rotated", lambda m: _incl(m, 1.57, 0.5, 0.04, 2.1).rotate(7.4),
     (48, 48)),
    ("inclined_conv_sheared", lambda m: m.Convolve([_incl(m, 0.78, 0.4, 0.1, 3.)
                                                    .shear(g1=0.1, g2=0.2),
                                                    m.Kolmogorov(fwhm=0.7)]), (48, 48)),
    ("tophat", lambda m: m.TopHat(radius=0.5, flux=3.), (40, 40)),
    ("tophat_sheared_conv", lambda m: m.Convolve([m.TopHat(radius=0.3, flux=2.)
                                                  .shear(g1=0.2, g2=0.1),
                                                  m.Moffat(3., fwhm=0.6)]), (40, 40)),
    ("delta_shifted_conv", lambda m: m.Convolve([m.DeltaFunction(flux=5.).shift(0.1, -0.15),
                                                 m.Moffat(2.5, fwhm=0.7)]), (40, 40)),
    ("inclined_sersic_15", lambda m: _incls(m, 1.5, 1.1, 0.1, 0.03, 4.), (48, 48)),
    ("inclined_sersic_trunc_rot", lambda m: _incls(m, 0.8, 1.4, 0.3, 0.03, 2., 1.2)
     .rotate(1.1), (48, 48)),
    ("inclined_sersic_conv", lambda m: m.Convolve([_incls(m, 2.5, 0.6, 0.05, 0.01, 5.),
                                                   m.Kolmogorov(fwhm=0.7)]), (48, 48)),
])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_profiles_parity(name, make, shape, precision):
    im, ref, info = _draw_pair(make, shape[0], shape[1], precision=precision)
    err = _peak_rel(im.array, ref)
    print(name, precision, "N", info['N'], "Nk", info['Nk'], "peak-rel", err)
    assert err < (F64_TOL if precision == 'f64' else F32_TOL)


def test_methods_and_centering():
    for method in ('fft', 'sb'):
        for utc in (True, False):
            make = lambda m: m.Convolve([m.Gaussian(sigma=0.5), m.Exponential(scale_radius=0.3)])
            prod, orc = _both(make)
            im = prod.drawImage(nx=30, ny=30, scale=0.25, dtype=np.float64, method=method,
                                use_true_center=utc)
            ref = orc.drawImage(30, 30, scale=0.25, method=method, use_true_center=utc)
            assert _peak_rel(im.array, ref) < F64_TOL


def test_auto_size():
    make = lambda m: m.Convolve([m.Exponential(half_light_radius=0.6), m.Moffat(3.5, fwhm=0.6)])
    prod, orc = _both(make)
    im = prod.drawImage(scale=0.2, dtype=np.float64)
    ref = orc.drawImage(scale=0.2)
    assert im.array.shape == ref.shape
    assert _peak_rel(im.array, ref) < F64_TOL


def test_golden_gauss_smallshear_convolve_box(golden_dir):
    # the reference's FFT-path golden (tests/test_convolve.py:220-236), 5 decimals
    saved = read_fits(os.path.join(golden_dir, "gauss_smallshear_convolve_box.fits"))
    psf = g.Gaussian(flux=1, sigma=1).shear(e1=0.04, e2=0.0)
    conv = g.Convolve([psf, g.Pixel(scale=0.2, flux=1.)])
    img = g.ImageF(saved.shape[1], saved.shape[0], scale=0.2)
    img.setCenter(0, 0)
    conv.drawImage(img, scale=0.2, method="sb", use_true_center=False)
    np.testing.assert_array_almost_equal(img.array, saved, 5)


def test_drawkimage_parity():
    make = lambda m: m.Exponential(scale_radius=13., flux=4.1)
    prod, orc = _both(make)
    for scale in (0.15 / 13., 0.6 / 13.):
        kim = g.ImageCD(45, 45, scale=scale)
        prod.drawKImage(kim)
        ref = orc.drawKImage(45, 45, scale)
        assert np.max(np.abs(kim.array - ref)) < 1e-12 * np.max(np.abs(ref))
    make2 = lambda m: m.Convolve([m.Sersic(2.2, half_light_radius=0.7).shear(g1=0.2, g2=0.3)
                                  .shift(0.3, 0.1), m.Kolmogorov(fwhm=0.5)])
    prod, orc = _both(make2)
    kim = g.ImageCD(31, 41, scale=0.3)
    prod.drawKImage(kim)
    ref = orc.drawKImage(31, 41, 0.3)
    assert np.max(np.abs(kim.array - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("Nk,N", [(64, 64), (96, 64), (130, 96), (256, 128)])
def test_wrap_and_irfft_kernels(Nk, N):
    rng = np.random.default_rng(Nk + N)
    karr = rng.normal(size=(Nk + 1, Nk // 2 + 1)) + 1j * rng.normal(size=(Nk + 1, Nk // 2 + 1))
    kw = render.wrap_hermx(karr, Nk, N)
    kw_ref = od.wrap_hermx(karr, Nk, N)
    np.testing.assert_allclose(kw, kw_ref, rtol=0, atol=1e-12)
    real = render.irfft(kw_ref, N)
    real_ref = od.irfft_shifted(kw_ref, N)
    assert np.max(np.abs(real - real_ref)) < 1e-13 * np.max(np.abs(real_ref))


def test_large_fft_sizes():
    # big Sersic galaxies need N = 1024 and beyond
    make = lambda m: m.Convolve([m.Sersic(4.5, half_light_radius=4.), m.Kolmogorov(fwhm=0.8)])
    im, ref, info = _draw_pair(make, 96, 96, precision='f64')
    print("N", info['N'], "peak-rel", _peak_rel(im.array, ref))
    assert info['N'] >= 512
    assert _peak_rel(im.array, ref) < F64_TOL


@pytest.mark.parametrize("precision,r0,moffat,N", [('f64', 22., True, 3072),
                                                   ('f64', 30., True, 4096),
                                                   ('f32', 75., False, 8192)])
def test_fft_sizes_beyond_64k_lds(precision, r0, moffat, N):
    # column transforms of 2 N complex values need more than 64 KiB of LDS here: 96 KiB and
    # 128 KiB in FP64, 128 KiB in FP32 (gfx950 allows 160 KiB per workgroup)
    def make(m):
        gal = m.Exponential(scale_radius=r0, flux=1e4).shear(g1=0.2, g2=0.1)
        return m.Convolve([gal, m.Moffat(beta=3, fwhm=0.7)]) if moffat else gal
    print("precision", precision, "expected N", N, flush=True)
    im, ref, info = _draw_pair(make, 64, 64, precision=precision)
    print("N", info['N'], "peak-rel", _peak_rel(im.array, ref))
    assert info['N'] == N
    assert _peak_rel(im.array, ref) < (F64_TOL if precision == 'f64' else F32_TOL)


@pytest.mark.parametrize("r0,moffat,N", [(40., True, 6144), (75., False, 8192)])
def test_fp64_fft_sizes_beyond_the_fused_kernels(r0, moffat, N):
    # N = 6144 and 8192 in FP64 need 192-256 KiB for one column transform: drawn as the
    # reference's separate steps (k-image, wrap, irfft with a four-step column transform)
    def make(m):
        gal = m.Exponential(scale_radius=r0, flux=1e4).shear(g1=0.2, g2=0.1)
        return m.Convolve([gal, m.Moffat(beta=3, fwhm=0.7)]) if moffat else gal
    print("expected N", N, flush=True)
    im, ref, info = _draw_pair(make, 64, 64, precision='f64')
    print("N", info['N'], "peak-rel", _peak_rel(im.array, ref))
    assert info['N'] == N
    assert _peak_rel(im.array, ref) < F64_TOL


@pytest.mark.parametrize("idx", range(4))
def test_drawkimage_phases_match_closed_form(idx):
    # the closed-form transformed Gaussians of tests/test_oracle_phases.py (ApplyKImagePhases'
    # separable and 2-d branches, offsets composed through outer Jacobians, flux scalings)
    # through the GPU's k-image program
    from tests.test_oracle_phases import _cases, closed_form
    name, obj, terms = _cases(g)[idx]
    nx, ny, dk = 41, 37, 0.21
    kim = g.ImageCD(nx, ny, scale=dk)
    obj.drawKImage(kim)
    assert kim.bounds.xmin == -(nx // 2) and kim.bounds.ymin == -(ny // 2)
    ref = closed_form(terms, nx, ny, dk)
    err = float(np.max(np.abs(kim.array - ref)) / np.max(np.abs(ref)))
    print(name, err)
    assert err < 1e-12, (name, err)


def test_golden_moffat_pixel(golden_dir):
    # the reference's FFT-path golden against an exact Maple calculation
    # (tests/test_convolve.py:29-60), 4 decimals, and the oracle at 1e-10
    saved = read_fits(os.path.join(golden_dir, "moffat_pixel.fits"))
    fw = 1.0927449310213702
    psf = g.Moffat(beta=1.5, fwhm=fw, trunc=4 * fw, flux=1)
    pixel = g.Pixel(scale=0.2, flux=1.)
    with pytest.warns(UserWarning, match="hard edges"):
        conv = g.Convolve([psf, pixel], real_space=False)
    img = g.ImageF(saved.shape[1], saved.shape[0], scale=0.2)
    img.setCenter(0, 0)
    conv.drawImage(img, scale=0.2, method="sb", use_true_center=False)
    np.testing.assert_array_almost_equal(img.array, saved, 4)
    ref = od.Convolution([od.Moffat(1.5, fwhm=fw, trunc=4 * fw, flux=1.),
                          od.Pixel(0.2, flux=1.)]).drawImage(
        saved.shape[1], saved.shape[0], scale=0.2, method='sb', use_true_center=False)
    img64 = g.ImageD(saved.shape[1], saved.shape[0], scale=0.2)
    img64.setCenter(0, 0)
    conv.drawImage(img64, scale=0.2, method="sb", use_true_center=False)
    assert _peak_rel(img64.array, ref) < F64_TOL
