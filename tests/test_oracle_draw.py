"""Pins for the oracle's drawImage(method='fft') chain (oracle/draw.py): the wrap closed form,
the inverse-FFT convention, and the reference's golden images."""
import os

import numpy as np
import pytest

from oracle import draw as od
from oracle import numerics as nm
from tests.fitsio import read_fits


@pytest.mark.parametrize("Nk,N", [(12, 12), (20, 12), (26, 8), (31 * 2, 16)])
def test_wrap_hermx_closed_form(Nk, N):
    # tests/test_image.py:3388-3453 in the reference: wrapping the implicitly Hermitian half
    # plane equals adding every point of the explicit full plane to its congruent target cell.
    rng = np.random.default_rng(Nk * 100 + N)
    half = Nk // 2
    full = rng.normal(size=(Nk + 1, Nk + 1)) + 1j * rng.normal(size=(Nk + 1, Nk + 1))
    full = 0.5 * (full + np.conj(full[::-1, ::-1]))     # full[j,i] == conj(full[-j,-i])
    stored = full[:, half:]                               # kx = 0..Nk/2, ky = -Nk/2..Nk/2
    got = od.wrap_hermx(stored, Nk, N)
    expect = np.zeros((N, N), dtype=complex)              # ii in [-N/2+1, N/2], jj in [-N/2, N/2-1]
    for j in range(-half, half + 1):
        for i in range(-half, half + 1):
            ii = (i - (-N // 2 + 1)) % N + (-N // 2 + 1)
            jj = (j - (-N // 2)) % N + (-N // 2)
            expect[jj + N // 2, ii + N // 2 - 1] += full[j + half, i + half]
    np.testing.assert_allclose(got, expect[:, N // 2 - 1:], rtol=0, atol=1e-12)


@pytest.mark.parametrize("N", [8, 12, 16])
def test_irfft_shifted_round_trip(N):
    # A real image on (-N/2..N/2-1)^2: forward DFT by brute force, then the reference's
    # irfft(shift_in=True, shift_out=True) convention must return the image.
    rng = np.random.default_rng(N)
    img = rng.normal(size=(N, N))
    x = np.arange(N) - N // 2
    kx = np.arange(N // 2 + 1)
    ky = np.arange(N) - N // 2
    ex = np.exp(-2j * np.pi * np.outer(kx, x) / N)         # (kx, x)
    ey = np.exp(-2j * np.pi * np.outer(ky, x) / N)         # (ky, y)
    W = ey @ img @ ex.T                                    # (ky, kx)
    np.testing.assert_allclose(od.irfft_shifted(W, N), img, rtol=0, atol=1e-12)


def _golden(golden_dir, name):
    return read_fits(os.path.join(golden_dir, name + ".fits"))


def test_golden_gauss_smallshear_convolve_box(golden_dir):
    # tests/test_convolve.py:220-236 in the reference (5 decimals): an FFT-path golden.
    saved = _golden(golden_dir, "gauss_smallshear_convolve_box")
    psf = od.Gaussian(flux=1, sigma=1).shear(e1=0.04, e2=0.0)
    conv = od.Convolve([psf, od.Pixel(0.2, flux=1.)])
    img = conv.drawImage(saved.shape[1], saved.shape[0], scale=0.2, method='sb',
                         use_true_center=False)
    np.testing.assert_array_almost_equal(img, saved, 5)


def _delta_conv(prof):
    # galsim_test_helpers.do_kvalue: convolve with a 1e-8 Gaussian to force the FFT path
    return od.Convolve([prof, od.Gaussian(sigma=1.e-8)])


@pytest.mark.parametrize("name,make,decimal", [
    ("gauss_1", lambda: od.Gaussian(flux=1, sigma=1), 5),
    ("kolmogorov", lambda: od.Kolmogorov(lam_over_r0=1.5, flux=1.8), 4),
])
def test_golden_real_space_images_via_fft(golden_dir, name, make, decimal):
    # Real-space 'sb' goldens (tests/test_gaussian.py:34, test_kolmogorov.py:44) drawn through
    # the FFT path, as the reference's do_kvalue check does (galsim_test_helpers.py:358-374,
    # 3 decimals there).  Cuspy profiles (exp_1.fits) cannot match a sampled cusp through a
    # band-limited FFT at these decimals and are pinned in k-space instead.
    saved = _golden(golden_dir, name)
    img = _delta_conv(make()).drawImage(saved.shape[1], saved.shape[0], scale=0.2,
                                        method='sb', use_true_center=False)
    np.testing.assert_array_almost_equal(img, saved, decimal)


def test_drawimage_sizes_and_flux():
    # drawFFT_makeKImage sizing: N >= 128, goodFFTSize, Nk > N only when maxk exceeds Nyquist
    gal = od.Exponential(half_light_radius=1.5, flux=1e4).shear(g1=0.1, g2=-0.05)
    psf = od.Moffat(beta=3, fwhm=0.7)
    img, info = od.Convolve([gal, psf]).drawImage(64, 64, scale=0.2, offset=(0.13, -0.21),
                                                   return_info=True)
    assert info['N'] == nm.good_fft_size(info['N']) and info['N'] >= 128
    assert info['Nk'] == info['N']
    # nearly all the flux lands in a 64x64 stamp of a 1.5" galaxy at 0.2"/pixel
    assert 0.9 * 1e4 < info['added_flux'] < 1e4
    assert abs(img.sum() - info['added_flux']) < 1e-6
    # the profile is centred on the (true) stamp centre plus the offset
    y, x = np.mgrid[0:64, 0:64]
    cx = (img * x).sum() / img.sum()
    cy = (img * y).sum() / img.sum()
    assert abs(cx - (31.5 + 0.13)) < 0.05 and abs(cy - (31.5 - 0.21)) < 0.05


def test_aliasing_fold_when_maxk_beyond_nyquist():
    # A small Kolmogorov has maxk > pi per pixel, so Nk > N and the k-image is folded.
    psf = od.Kolmogorov(fwhm=0.6)
    img, info = psf.drawImage(32, 32, scale=0.2, return_info=True)
    assert info['Nk'] > info['N']
    # folding preserves the total flux in the periodic image (the k=0 sample is untouched)
    assert abs(img.sum() - 1.) < 5e-3


def test_golden_moffat_pixel(golden_dir):
    # the reference's FFT-path golden against an exact Maple calculation
    # (tests/test_convolve.py:29-60): a truncated Moffat(beta=1.5) convolved with a Pixel by FFT,
    # drawn method='sb', use_true_center=False, at 4 decimals
    saved = read_fits(os.path.join(golden_dir, "moffat_pixel.fits"))
    fw = 1.0927449310213702
    conv = od.Convolution([od.Moffat(1.5, fwhm=fw, trunc=4 * fw, flux=1.),
                           od.Pixel(0.2, flux=1.)])
    ny, nx = saved.shape
    img = conv.drawImage(nx, ny, scale=0.2, method='sb', use_true_center=False)
    np.testing.assert_array_almost_equal(img.astype(np.float32), saved, 4)
