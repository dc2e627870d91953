"""GPU parity of galsim_amd.fft (fft2 / ifft2 / rfft2 / irfft2) and Image.calculate_fft /
calculate_inverse_fft against the oracle (numpy.fft with the reference's shift conventions,
pinned in tests/test_fft_cpu.py).  Tolerance: the reference's own test asks for 9 decimals
against numpy (tests/test_draw.py:1343-1521); here every result is also within 1e-12 of the
largest output value."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import galsim_amd as g
from oracle import fft as of

SHIFTS = [(False, False), (True, False), (False, True), (True, True)]
REL = 1e-12


def _close(got, ref, rel=REL):
    got = np.asarray(got)
    assert got.shape == ref.shape
    np.testing.assert_almost_equal(got, ref, 9)
    scale = max(np.max(np.abs(ref)), 1e-300)
    assert np.max(np.abs(got - ref)) <= rel * scale, np.max(np.abs(got - ref)) / scale


def _inputs():
    # the input list of tests/test_draw.py:1347-1378 (noise from numpy instead of galsim)
    rng = np.random.default_rng(1234)
    out = [np.array([[0, 1, 2, 1], [1, 2, 3, 2], [2, 3, 4, 3], [1, 2, 3, 2]], dtype=int),
           np.array([[0, 1], [1, 2], [2, 3], [1, 2]], dtype=int)]
    sq = [(n, n) for n in (2, 4, 8, 10)]
    rect = [(4, 2), (2, 4), (6, 10), (10, 6)]
    for s in sq + rect:
        out.append(5. * rng.normal(size=s))
    for s in sq + rect:
        out.append(5. * rng.normal(size=s) + 5j * rng.normal(size=s))
    return out


@pytest.mark.parametrize("si,so", SHIFTS)
def test_matches_numpy_on_reference_inputs(si, so):
    for xar in _inputs():
        print("shape", xar.shape, xar.dtype, "shifts", si, so, flush=True)
        kar = g.fft.fft2(xar, si, so)
        _close(kar, of.fft2(xar, si, so))
        _close(g.fft.ifft2(kar, so, si), of.ifft2(kar, so, si))
        _close(g.fft.ifft2(kar, so, si), xar.astype(np.complex128))
        if np.iscomplexobj(xar):
            continue
        rk = g.fft.rfft2(xar, si, so)
        _close(rk, of.rfft2(xar, si, so))
        _close(g.fft.irfft2(rk, so, si), of.irfft2(rk, so, si))
        _close(g.fft.irfft2(rk, so, si), xar.astype(np.float64))
        # ifft2 takes real arrays too (tests/test_draw.py:1477-1486)
        _close(g.fft.fft2(g.fft.ifft2(xar, si, so), so, si), xar.astype(np.complex128))


# direct mixed-radix lengths (2, 3, 5 factors) and Bluestein lengths (prime factors 7, 11, 23,
# 131, 587)
SIZES = [(64, 96), (128, 128), (250, 384), (14, 22), (46, 262), (1174, 1174), (2048, 1536),
         (300, 2048), (6144, 40), (8192, 64),
         # every power-of-two line length of the register-resident transforms (128..4096) along
         # rows and columns, line counts below and off a multiple of a workgroup's lines
         (256, 512), (1024, 1024), (4096, 256), (512, 4096), (6, 256), (4096, 4), (2048, 2048),
         # and the 3 x 2^k lengths (192..3072), twelve values per thread
         (1536, 768), (384, 3072), (192, 1536), (3072, 384), (768, 3072)]


@pytest.mark.parametrize("shape", SIZES)
def test_sizes_against_numpy(shape):
    print("shape", shape, flush=True)
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    x = rng.normal(size=shape)
    z = x + 1j * rng.normal(size=shape)
    for si, so in ((False, False), (True, True)):
        _close(g.fft.fft2(z, si, so), of.fft2(z, si, so))
        _close(g.fft.ifft2(z, si, so), of.ifft2(z, si, so))
        rk = g.fft.rfft2(x, si, so)
        _close(rk, of.rfft2(x, si, so))
        _close(g.fft.fft2(x, si, so), of.fft2(x, si, so))
        _close(g.fft.irfft2(rk, so, si), x)


@pytest.mark.parametrize("n", [8192, 6144])
def test_long_real_rows(n):
    # rfft2 / irfft2 rows of 8192 (6144) run as complex transforms of 4096 (3072); complex rows
    # that long exceed one workgroup's LDS and raise
    rng = np.random.default_rng(n)
    x = rng.normal(size=(64, n))
    for si, so in ((False, False), (True, True)):
        rk = g.fft.rfft2(x, si, so)
        _close(rk, of.rfft2(x, si, so))
        _close(g.fft.irfft2(rk, so, si), x)


def test_unaligned_buffers_take_the_lds_kernels():
    # the register-resident transforms need 16-byte aligned buffers: a float64 buffer starting
    # 8 bytes into an allocation takes the LDS kernels, with the same results
    import torch
    from galsim_amd import _lib
    L = _lib.lib()
    M, N = 256, 512
    W = N // 2 + 1
    rng = np.random.default_rng(16)
    x = rng.normal(size=(M, N))
    stream = torch.cuda.current_stream().cuda_stream
    xin = torch.zeros(M * N + 1, dtype=torch.float64, device='cuda')
    xin[1:] = torch.from_numpy(x.ravel()).cuda()
    assert xin.data_ptr() % 16 == 0
    k = torch.empty((M, W), dtype=torch.complex128, device='cuda')
    tmp = torch.empty((M, W), dtype=torch.complex128, device='cuda')
    for si, so in ((False, False), (True, True)):
        print("unaligned rfft2/irfft2", M, N, si, so, flush=True)
        _lib.check(L.gsmi_rfft2_f64(xin.data_ptr() + 8, M, N, int(si), int(so), k.data_ptr(),
                                    tmp.data_ptr(), stream), "gsmi_rfft2_f64")
        _close(k.cpu().numpy(), of.rfft2(x, si, so))
        xout = torch.zeros(M * N + 1, dtype=torch.float64, device='cuda')
        _lib.check(L.gsmi_irfft2_f64(k.data_ptr(), M, N, int(so), int(si),
                                     xout.data_ptr() + 8, tmp.data_ptr(), stream),
                   "gsmi_irfft2_f64")
        _close(xout[1:].reshape(M, N).cpu().numpy(), x)


def test_column_passes_query():
    from galsim_amd import _lib
    L = _lib.lib()
    assert L.gsmi_fft_column_passes(2048) == 1
    assert L.gsmi_fft_column_passes(8192) == 2
    assert L.gsmi_fft_column_passes(7) < 0


def test_device_tensors_stay_on_device():
    import torch
    rng = np.random.default_rng(5)
    x = rng.normal(size=(96, 128))
    t = torch.from_numpy(x).cuda()
    k = g.fft.rfft2(t, True, True)
    assert isinstance(k, torch.Tensor) and k.is_cuda and k.dtype == torch.complex128
    _close(k.cpu().numpy(), of.rfft2(x, True, True))
    back = g.fft.irfft2(k, True, True)
    assert back.is_cuda and back.dtype == torch.float64
    _close(back.cpu().numpy(), x)


def test_image_fft_round_trips():
    # tests/test_draw.py:1214-1272 in the reference, for each input dtype
    for dt in (np.int16, np.int32, np.float32, np.float64, np.complex128, int, float, complex):
        print("dtype", dt, flush=True)
        a = np.array([[0, 2, 4, 2], [2, 4, 6, 4], [4, 6, 8, 4], [2, 4, 6, 6]]).astype(dt)
        xim = g.Image(a, xmin=-2, ymin=-2)
        xim.scale = 0.1
        kim = xim.calculate_fft()
        kref, dk = of.calculate_fft(a.astype(np.complex128), -2, -2, 0.1)
        _close(kim.array, kref)
        assert kim.bounds == g.BoundsI(0, 2, -2, 1) and kim.scale == pytest.approx(dk, rel=1e-15)
        xim2 = kim.calculate_inverse_fft()
        np.testing.assert_almost_equal(xim2.array, xim.array.real)
        assert xim2.bounds == g.BoundsI(-2, 1, -2, 1)

        k = np.array([[4, 2, 0], [6, 4, 2], [8, 6, 4], [6, 4, 2]]).astype(dt)
        kim = g.Image(k, xmin=0, ymin=-2)
        kim.scale = 0.1
        x = kim.calculate_inverse_fft()
        xref, dx = of.calculate_inverse_fft(k, 0, -2, 0.1)
        _close(x.array, xref)
        np.testing.assert_almost_equal(x.calculate_fft().array, kim.array)

        k3 = np.array([[0, 1, 2, 1, 0], [1, 4, 6, 4, 1], [2, 6, 8, 6, 2], [1, 4, 6, 4, 1],
                       [0, 1, 2, 1, 0]]).astype(dt)
        kim3 = g.Image(k3, xmin=-2, ymin=-2)
        kim3.scale = 0.1
        x = kim3.calculate_inverse_fft()
        _close(x.array, of.calculate_inverse_fft(k3, -2, -2, 0.1)[0])
        np.testing.assert_almost_equal(x.calculate_fft().array, kim.array)

        xa = g.Image(np.array([[0, 0, 0, 0], [2, 4, 6, 0], [4, 6, 8, 0],
                               [0, 0, 0, 0]]).astype(dt), xmin=-2, ymin=-2)
        xb = g.Image(np.array([[2, 4, 6], [4, 6, 8]]).astype(dt), xmin=-2, ymin=-1)
        xa.scale = xb.scale = 0.1
        np.testing.assert_almost_equal(xa.calculate_fft().array, xb.calculate_fft().array)

        ka = g.Image(np.array([[4, 2, 0], [6, 4, 0], [8, 6, 0], [6, 4, 0]]).astype(dt),
                     xmin=0, ymin=-2)
        kb = g.Image(np.array([[6, 4], [8, 6], [6, 4], [4, 2]]).astype(dt), xmin=0, ymin=-1)
        ka.scale = kb.scale = 0.1
        np.testing.assert_almost_equal(ka.calculate_inverse_fft().array,
                                       kb.calculate_inverse_fft().array)


def test_drawkimage_inverse_fft_1174():
    # tests/test_draw.py:1274-1321 in the reference: Moffat(beta=1.5) with maxk_threshold=1e-4
    # gives a 1175 x 1175 k-image (N = 1174, N/2 odd and a prime 587) and, at N = 1536, a size
    # that goes straight through the mixed-radix transforms.
    obj = g.Moffat(flux=1.7, beta=1.5, scale_radius=0.5).withGSParams(maxk_threshold=1.e-4)
    im1 = obj.drawKImage()
    assert im1.bounds == g.BoundsI(-587, 587, -587, 587)
    real = im1.calculate_inverse_fft()
    ref, dx = of.calculate_inverse_fft(im1.array, -587, -587, im1.scale)
    assert real.bounds == g.BoundsI(-587, 586, -587, 586)
    assert real.scale == pytest.approx(dx, rel=1e-15)
    _close(real.array, ref)
    N = 1536
    im2 = obj.drawKImage(nx=N + 1, ny=N + 1, scale=2. * np.pi / (N * obj.nyquist_scale))
    real2 = im2.calculate_inverse_fft()
    _close(real2.array, of.calculate_inverse_fft(im2.array, -N // 2, -N // 2, im2.scale)[0])
    # and back: the forward transform of the real image
    k2 = real2.calculate_fft()
    kref, _ = of.calculate_fft(real2.array, -N // 2, -N // 2, real2.scale)
    _close(k2.array, kref)
