"""Oracle for galsim.fft and Image.calculate_fft / calculate_inverse_fft.

Test infrastructure only (see ``oracle/__init__.py``).

Restates, in numpy FP64:
* fft2 / ifft2 / rfft2 / irfft2 with shift_in / shift_out (galsim/fft.py:29-300) as numpy.fft with
  numpy.fft.fftshift applied to the input / output, which is what the reference's docstrings state
  and tests/test_draw.py:1343-1521 checks; fft2 of a real array is the rfft2 half-plane plus the
  Hermitian fill of galsim/fft.py:88-103;
* Image.calculate_fft (galsim/image.py:915-956) and calculate_inverse_fft (galsim/image.py:958-1016)
  with wrapImage(hermx=True) from ``oracle.draw``.
The numpy conventions are pinned by brute-force DFT sums in tests/test_fft_cpu.py.
"""
import numpy as np

from . import draw as od


def _check(a, real_half=False):
    s = np.shape(a)
    if len(s) != 2:
        raise ValueError("Input array must be 2D.", s)
    if s[0] % 2 or (not real_half and s[1] % 2):
        raise ValueError("Input array must have even sizes.", s)
    return s


def fft2(a, shift_in=False, shift_out=False):
    """galsim/fft.py:29-106."""
    M, N = _check(a)
    a = np.asarray(a)
    if a.dtype.kind == 'c':
        x = np.fft.fftshift(a) if shift_in else a
        k = np.fft.fft2(x.astype(np.complex128))
        return np.fft.fftshift(k) if shift_out else k
    rk = rfft2(a, shift_in, shift_out)
    Mo2, No2 = M // 2, N // 2
    kar = np.empty((M, N), dtype=np.complex128)
    # the fill of galsim/fft.py:88-103
    if shift_out:
        kar[:, No2:N] = rk[:, 0:No2]
        kar[0, 0:No2] = rk[0, No2:0:-1].conjugate()
        kar[1:Mo2, 0:No2] = rk[M - 1:Mo2:-1, No2:0:-1].conjugate()
        kar[Mo2:M, 0:No2] = rk[Mo2:0:-1, No2:0:-1].conjugate()
    else:
        kar[:, 0:No2] = rk[:, 0:No2]
        kar[0, No2:N] = rk[0, No2:0:-1].conjugate()
        kar[1:M, No2:N] = rk[M - 1:0:-1, No2:0:-1].conjugate()
    return kar


def ifft2(a, shift_in=False, shift_out=False):
    """galsim/fft.py:109-174."""
    _check(a)
    a = np.asarray(a)
    x = np.fft.fftshift(a) if shift_in else a
    out = np.fft.ifft2(x.astype(np.complex128))
    return np.fft.fftshift(out) if shift_out else out


def rfft2(a, shift_in=False, shift_out=False):
    """galsim/fft.py:177-230 (a complex input contributes its real part, as src/Image.cpp:761-788
    reads it)."""
    _check(a)
    a = np.asarray(a)
    a = (a.real if a.dtype.kind == 'c' else a).astype(np.float64)
    x = np.fft.fftshift(a) if shift_in else a
    k = np.fft.rfft2(x)
    return np.fft.fftshift(k, axes=(0,)) if shift_out else k


def irfft2(a, shift_in=False, shift_out=False):
    """galsim/fft.py:233-300: an M x (N/2+1) half-plane to an M x N real image."""
    M, W = _check(a, real_half=True)
    N = 2 * (W - 1)
    a = np.asarray(a).astype(np.complex128)
    x = np.fft.fftshift(a, axes=(0,)) if shift_in else a
    out = np.fft.irfft2(x, s=(M, N))
    return np.fft.fftshift(out) if shift_out else out


def calculate_fft(arr, xmin, ymin, dx):
    """Image.calculate_fft (galsim/image.py:915-956) of the image ``arr`` with origin
    (xmin, ymin) and pixel scale dx -> (karr on (0..No2, -No2..No2-1), dk)."""
    ny, nx = arr.shape
    xmax, ymax = xmin + nx - 1, ymin + ny - 1
    No2 = max(-xmin, xmax + 1, -ymin, ymax + 1)
    full = np.zeros((2 * No2, 2 * No2))
    a = arr.real if np.iscomplexobj(arr) else arr
    full[ymin + No2: ymax + No2 + 1, xmin + No2: xmax + No2 + 1] = a
    k = rfft2(full, shift_in=True, shift_out=True) * (dx * dx)
    return k, np.pi / (No2 * dx)


def calculate_inverse_fft(karr, xmin, ymin, dk):
    """Image.calculate_inverse_fft (galsim/image.py:958-1016) of the k-image ``karr`` with origin
    (xmin, ymin) and scale dk -> (real image on (-No2..No2-1)^2, dx)."""
    ny, nx = karr.shape
    xmax, ymax = xmin + nx - 1, ymin + ny - 1
    if not (xmin <= 0 <= xmax and ymin <= 0 <= ymax):
        raise ValueError("calculate_inverse_fft requires that the image includes (0,0)")
    No2 = max(xmax, -ymin, ymax)
    N = 2 * No2
    a = np.asarray(karr).astype(np.complex128)
    if (xmin, xmax, ymin, ymax) == (0, No2, -No2, No2 - 1):
        kw = a
    else:
        full = np.zeros((N + 1, No2 + 1), dtype=np.complex128)
        full[ymin + No2: ymax + No2 + 1, 0: xmax + 1] = a[:, -xmin:]
        kw = od.wrap_hermx(full, N, N)
    real = irfft2(kw, shift_in=True, shift_out=True) * (dk * No2 / np.pi) ** 2
    return real, np.pi / (No2 * dk)
