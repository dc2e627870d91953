"""galsim.fft on the GPU: fft2, ifft2, rfft2 and irfft2 of 2-d arrays with even sizes, with the
reference's shift_in / shift_out options (galsim/fft.py:29-300).

Each function takes a numpy array (returned as numpy, like the reference) or a torch tensor on the
GPU (returned as a tensor on the GPU, without a host round trip).  The transforms run in libgsmi
(gsmi_cfft2_f64, gsmi_rfft2_f64, gsmi_irfft2_f64, gsmi_hermitian_fill_f64) in float64 /
complex128; results equal numpy.fft's to rounding.
"""
import numpy as np

from . import _lib
from ._lib import GalSimAMDValueError

__all__ = ['fft2', 'ifft2', 'rfft2', 'irfft2']


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.GalSimAMDError("galsim_amd.fft runs on the GPU; no HIP device is available")
    return torch


def _as_device(a, complex_in):
    """(device tensor, was_device): the input as a contiguous float64 / complex128 GPU tensor."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a
        dev = True
    else:
        arr = np.asarray(a)
        t = torch.from_numpy(np.ascontiguousarray(arr))
        dev = False
    if complex_in:
        t = t.to(torch.complex128) if t.is_complex() else t.to(torch.float64).to(torch.complex128)
    else:
        t = t.real if t.is_complex() else t
        t = t.to(torch.float64)
    return t.to('cuda').contiguous(), dev


def _shape(a):
    s = tuple(a.shape)
    if len(s) != 2:
        raise GalSimAMDValueError("Input array must be 2D.", s)
    return s


def _even(s):
    M, N = s
    if M % 2 or N % 2:
        raise GalSimAMDValueError("Input array must have even sizes.", s)
    return M, N


def _is_complex(a):
    try:
        import torch
        if isinstance(a, torch.Tensor):
            return a.is_complex()
    except ImportError:
        pass
    return np.asarray(a).dtype.kind == 'c'


def _out(t, dev):
    return t if dev else t.cpu().numpy()


def _stream(torch):
    return torch.cuda.current_stream().cuda_stream


def _cfft(a, inverse, shift_in, shift_out):
    M, N = _even(_shape(a))
    x, dev = _as_device(a, True)
    torch = _torch()
    out = torch.empty((M, N), dtype=torch.complex128, device='cuda')
    tmp = torch.empty((M, N), dtype=torch.complex128, device='cuda')
    _lib.check(_lib.lib().gsmi_cfft2_f64(x.data_ptr(), M, N, int(bool(inverse)),
                                         int(bool(shift_in)), int(bool(shift_out)),
                                         out.data_ptr(), tmp.data_ptr(), _stream(torch)),
               "gsmi_cfft2_f64")
    return out, dev


def _rfft(x, M, N, shift_in, shift_out):
    torch = _torch()
    W = N // 2 + 1
    out = torch.empty((M, W), dtype=torch.complex128, device='cuda')
    tmp = torch.empty((M, W), dtype=torch.complex128, device='cuda')
    _lib.check(_lib.lib().gsmi_rfft2_f64(x.data_ptr(), M, N, int(bool(shift_in)),
                                         int(bool(shift_out)), out.data_ptr(), tmp.data_ptr(),
                                         _stream(torch)), "gsmi_rfft2_f64")
    return out


def irfft_scratch_rows(M):
    """Rows of (N/2+1) complex scratch gsmi_irfft2_f64 needs: M, or 2 M when columns of length
    M take the four-step transform on this device."""
    passes = _lib.lib().gsmi_fft_column_passes(int(M))
    if passes < 0:
        _lib.check(passes, "gsmi_fft_column_passes")
    return 2 * M if passes == 2 else M


def fft2(a, shift_in=False, shift_out=False):
    """numpy.fft.fft2(a) for a 2-d array with even sizes (galsim/fft.py:29-106): complex input is
    transformed as complex128; real input through rfft2 and the Hermitian fill of the kx < 0 half,
    as the reference does.  shift_in / shift_out apply numpy.fft.fftshift to the input / output."""
    M, N = _even(_shape(a))
    if _is_complex(a):
        out, dev = _cfft(a, False, shift_in, shift_out)
        return _out(out, dev)
    x, dev = _as_device(a, False)
    torch = _torch()
    half = _rfft(x, M, N, shift_in, shift_out)
    out = torch.empty((M, N), dtype=torch.complex128, device='cuda')
    _lib.check(_lib.lib().gsmi_hermitian_fill_f64(half.data_ptr(), M, N, int(bool(shift_out)),
                                                  out.data_ptr(), _stream(torch)),
               "gsmi_hermitian_fill_f64")
    return _out(out, dev)


def ifft2(a, shift_in=False, shift_out=False):
    """numpy.fft.ifft2(a) for a 2-d array with even sizes (galsim/fft.py:109-174); real input is
    taken as complex with zero imaginary part."""
    out, dev = _cfft(a, True, shift_in, shift_out)
    return _out(out, dev)


def rfft2(a, shift_in=False, shift_out=False):
    """numpy.fft.rfft2(a) for a real 2-d array with even sizes (galsim/fft.py:177-230): an
    M x (N/2+1) complex128 result; shift_out applies numpy.fft.fftshift along axis 0 only."""
    M, N = _even(_shape(a))
    x, dev = _as_device(a, False)
    return _out(_rfft(x, M, N, shift_in, shift_out), dev)


def irfft2(a, shift_in=False, shift_out=False):
    """numpy.fft.irfft2(a) for an M x (N/2+1) array, M even, giving an M x N real image
    (galsim/fft.py:233-300); shift_in applies numpy.fft.fftshift along axis 0 of the input,
    shift_out numpy.fft.fftshift to the output."""
    s = _shape(a)
    M, No2 = s[0], s[1] - 1
    if M % 2:
        raise GalSimAMDValueError("Input array must have even sizes.", s)
    if No2 < 1:
        raise GalSimAMDValueError("Input array must have at least 2 columns.", s)
    N = 2 * No2
    k, dev = _as_device(a, True)
    torch = _torch()
    out = torch.empty((M, N), dtype=torch.float64, device='cuda')
    tmp = torch.empty((irfft_scratch_rows(M), No2 + 1), dtype=torch.complex128, device='cuda')
    _lib.check(_lib.lib().gsmi_irfft2_f64(k.data_ptr(), M, N, int(bool(shift_in)),
                                          int(bool(shift_out)), out.data_ptr(), tmp.data_ptr(),
                                          _stream(torch)), "gsmi_irfft2_f64")
    return _out(out, dev)
