"""gsmi_irfft_f64_noalloc: the allocation-free irfft entry point (static device scratch) against
gsmi_irfft_f64 (caller scratch) and the oracle's irfft, on consecutive calls of falling size."""
import ctypes

import numpy as np
import pytest

from oracle import draw as od

pytestmark = pytest.mark.gpu


def _calls(kw_host, N):
    import torch
    from galsim_amd import _lib
    lib = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    k = torch.from_numpy(kw_host.view(np.float64).copy()).cuda()
    out_ref = torch.full((N, N), float('nan'), dtype=torch.float64, device='cuda')
    scratch = torch.empty(N * (N // 2 + 1) * 2, dtype=torch.float64, device='cuda')
    _lib.check(lib.gsmi_irfft_f64(k.data_ptr(), N, out_ref.data_ptr(), scratch.data_ptr(),
                                  stream), "gsmi_irfft_f64")
    out = torch.full((N, N), float('nan'), dtype=torch.float64, device='cuda')
    _lib.check(lib.gsmi_irfft_f64_noalloc(k.data_ptr(), N, out.data_ptr(), stream),
               "gsmi_irfft_f64_noalloc")
    torch.cuda.synchronize()
    return out.cpu().numpy(), out_ref.cpu().numpy()


def _kimage(N, seed):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.normal(size=(N, N // 2 + 1))
                                + 1j * rng.normal(size=(N, N // 2 + 1)))


@pytest.mark.parametrize("sizes", [(1536, 256), (2048, 192, 128)])
def test_noalloc_irfft_matches_allocating_entry(sizes):
    for i, N in enumerate(sizes):
        print("noalloc irfft call", i, "N", N, flush=True)
        kw = _kimage(N, 100 * i + N)
        got, ref = _calls(kw, N)
        assert np.all(np.isfinite(got))
        # same kernels, same twiddle values: identical results
        np.testing.assert_array_equal(got, ref)
        real_ref = od.irfft_shifted(kw, N)
        assert np.max(np.abs(got - real_ref)) < 1e-13 * np.max(np.abs(real_ref))


def test_noalloc_irfft_rejects_sizes_beyond_its_scratch():
    from galsim_amd import _lib
    lib = _lib.lib()
    # checked on the host before any device work: a null-free call with N = 4096 is refused
    dummy = ctypes.c_double(0.)
    rc = lib.gsmi_irfft_f64_noalloc(ctypes.addressof(dummy), 4096, ctypes.addressof(dummy), None)
    assert rc != 0
    assert b"4096" in lib.gsmi_last_error()
    rc = lib.gsmi_irfft_f64_noalloc(ctypes.addressof(dummy), 1000, ctypes.addressof(dummy), None)
    assert rc != 0
