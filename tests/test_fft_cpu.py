"""CPU pins for the galsim.fft oracle (oracle/fft.py) and the host-side argument checks of
galsim_amd.fft and Image.calculate_fft / calculate_inverse_fft.

The oracle states the shift options as numpy.fft.fftshift of the input / output.  The reference's
C++ (src/Image.cpp:729-990) implements them instead with alternating signs on the input and
output of FFTW's transforms; the brute-force transforms below restate those sign loops over
explicit DFT sums, so a wrong axis, direction or normalisation in the oracle fails here."""
import numpy as np
import pytest

import galsim_amd as g
from oracle import fft as of


def _dft2(x, sign):
    """sum_{j,i} x[j, i] exp(sign 2 pi i (ky j / M + kx i / N)), by explicit matrices."""
    M, N = x.shape
    FM = np.exp(sign * 2j * np.pi * np.outer(np.arange(M), np.arange(M)) / M)
    FN = np.exp(sign * 2j * np.pi * np.outer(np.arange(N), np.arange(N)) / N)
    return FM @ x @ FN


def _checker(M, N):
    return (-1.) ** np.add.outer(np.arange(M), np.arange(N))


def cfft_signs(x, inverse, shift_in, shift_out):
    """cfft of src/Image.cpp:913-990: input scaled (and sign-alternated for shift_out), FFTW
    transform, output sign-alternated for shift_in."""
    M, N = x.shape
    fac = 1. / (M * N) if inverse else 1.
    a = x.astype(np.complex128) * fac
    if shift_out:
        s0 = -1. if (shift_in and (N // 2 + M // 2) % 2 == 1) else 1.
        a = a * s0 * _checker(M, N)
    k = _dft2(a, +1 if inverse else -1)
    if shift_in:
        k = k * _checker(M, N)
    return k


def rfft_signs(x, shift_in, shift_out):
    """rfft of src/Image.cpp:729-812: rows sign-alternated for shift_out (starting at -1 when
    shift_in and Ny/2 is odd), FFTW r2c, kx >= 0 half sign-alternated for shift_in."""
    M, N = x.shape
    a = np.asarray(x, dtype=np.float64).copy()
    if shift_out:
        s0 = -1. if (shift_in and (M // 2) % 2 == 1) else 1.
        a *= s0 * ((-1.) ** np.arange(M))[:, None]
    k = _dft2(a, -1)[:, :N // 2 + 1]
    if shift_in:
        k = k * _checker(M, N // 2 + 1)
    return k


def irfft_signs(h, shift_in, shift_out):
    """irfft of src/Image.cpp:814-911: rows taken from Ny/2 on for shift_in, scaled by 1/(Nx Ny)
    and sign-alternated for shift_out, then the real inverse (kx = 0 and N/2 real parts)."""
    M, W = h.shape
    N = 2 * (W - 1)
    a = np.asarray(h, dtype=np.complex128)
    if shift_in:
        a = np.concatenate([a[M // 2:], a[:M // 2]])
    a = a / (M * N)
    if shift_out:
        a = a * _checker(M, W)
    # unnormalised complex inverse along y, then the real inverse along x
    ay = np.exp(2j * np.pi * np.outer(np.arange(M), np.arange(M)) / M) @ a
    i = np.arange(N)
    out = np.outer(ay[:, 0].real, np.ones(N)) + np.outer(ay[:, W - 1].real, (-1.) ** i)
    for kx in range(1, W - 1):
        out += 2. * (ay[:, kx][:, None] * np.exp(2j * np.pi * kx * i / N)[None, :]).real
    return out


SHAPES = [(2, 2), (4, 4), (2, 4), (4, 2), (6, 10), (10, 6), (8, 8), (10, 10)]
SHIFTS = [(False, False), (True, False), (False, True), (True, True)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("si,so", SHIFTS)
def test_oracle_complex_matches_sign_form(shape, si, so):
    rng = np.random.default_rng(hash(shape) % 1000)
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    np.testing.assert_allclose(of.fft2(x, si, so), cfft_signs(x, False, si, so), atol=1e-12)
    np.testing.assert_allclose(of.ifft2(x, si, so), cfft_signs(x, True, si, so), atol=1e-13)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("si,so", SHIFTS)
def test_oracle_real_matches_sign_form(shape, si, so):
    rng = np.random.default_rng(hash(shape) % 1000 + 7)
    x = rng.normal(size=shape)
    np.testing.assert_allclose(of.rfft2(x, si, so), rfft_signs(x, si, so), atol=1e-12)
    # fft2 of a real array (rfft2 + Hermitian fill) is the full complex transform
    np.testing.assert_allclose(of.fft2(x, si, so), cfft_signs(x, False, si, so), atol=1e-12)
    # irfft2 of a Hermitian half-plane
    h = of.rfft2(x, si, so)
    np.testing.assert_allclose(of.irfft2(h, so, si), irfft_signs(h, so, si), atol=1e-13)
    np.testing.assert_allclose(of.irfft2(h, so, si), x, atol=1e-13)


def test_oracle_image_round_trips():
    # the hand-made images of tests/test_draw.py:1214-1272 in the reference
    xim = np.array([[0, 2, 4, 2], [2, 4, 6, 4], [4, 6, 8, 4], [2, 4, 6, 6]], float)
    k, dk = of.calculate_fft(xim, -2, -2, 0.1)
    x2, dx = of.calculate_inverse_fft(k, 0, -2, dk)
    np.testing.assert_allclose(x2, xim, atol=1e-12)
    assert dx == pytest.approx(0.1, rel=1e-15)
    kim = np.array([[4, 2, 0], [6, 4, 2], [8, 6, 4], [6, 4, 2]], float)
    x, dx = of.calculate_inverse_fft(kim, 0, -2, 0.1)
    k2, _ = of.calculate_fft(x, -2, -2, dx)
    np.testing.assert_allclose(k2, kim, atol=1e-12)
    kim3 = np.array([[0, 1, 2, 1, 0], [1, 4, 6, 4, 1], [2, 6, 8, 6, 2], [1, 4, 6, 4, 1],
                     [0, 1, 2, 1, 0]], float)
    x, dx = of.calculate_inverse_fft(kim3, -2, -2, 0.1)
    k2, _ = of.calculate_fft(x, -2, -2, dx)
    np.testing.assert_allclose(k2, kim, atol=1e-12)
    # zero padding of the x image
    xa = np.array([[0, 0, 0, 0], [2, 4, 6, 0], [4, 6, 8, 0], [0, 0, 0, 0]], float)
    xb = np.array([[2, 4, 6], [4, 6, 8]], float)
    np.testing.assert_allclose(of.calculate_fft(xa, -2, -2, 0.1)[0],
                               of.calculate_fft(xb, -2, -1, 0.1)[0], atol=1e-12)
    # zero padding of the k image
    ka = np.array([[4, 2, 0], [6, 4, 0], [8, 6, 0], [6, 4, 0]], float)
    kb = np.array([[6, 4], [8, 6], [6, 4], [4, 2]], float)
    np.testing.assert_allclose(of.calculate_inverse_fft(ka, 0, -2, 0.1)[0],
                               of.calculate_inverse_fft(kb, 0, -1, 0.1)[0], atol=1e-12)


def test_oracle_planted_mistake_is_caught():
    # a shift applied along the wrong axis must fail the sign-form comparison
    rng = np.random.default_rng(3)
    x = rng.normal(size=(6, 10))
    wrong = np.fft.fftshift(np.fft.rfft2(x), axes=(1,))
    assert np.max(np.abs(wrong - rfft_signs(x, False, True))) > 1e-3


def test_argument_checks_without_gpu():
    # tests/test_draw.py:1488-1521 in the reference
    a = np.arange(16.).reshape(4, 4)
    for f in (g.fft.fft2, g.fft.ifft2, g.fft.rfft2, g.fft.irfft2):
        for bad in (a.ravel(), a.reshape(2, 2, 4), a.reshape(2, 2, 2, 2)):
            with pytest.raises(ValueError):
                f(bad)
        for bad in (a[:3, :3], a[:3, :]):
            with pytest.raises(ValueError):
                f(bad)
    for f in (g.fft.fft2, g.fft.ifft2, g.fft.rfft2):
        with pytest.raises(ValueError):
            f(a[:, :3])
    im = g.Image(a, xmin=-2, ymin=-2)
    with pytest.raises(g.GalSimError):
        im.calculate_fft()                  # no scale
    with pytest.raises(g.GalSimError):
        im.calculate_inverse_fft()
    im.scale = 1.
    im.setOrigin(1, 1)
    with pytest.raises(g.GalSimAMDBoundsError):
        im.calculate_inverse_fft()          # (0,0) not included
    with pytest.raises(g.GalSimAMDUndefinedBoundsError):
        g.Image(bounds=g.BoundsI(), scale=1.).calculate_fft()
