"""The ``_galsim`` module's interface for GalSim's FFT drawing path, on the GPU.

GalSim's Python layer reaches its C++ core through the pybind11 module ``_galsim``
(pysrc/module.cpp:72-111).  For ``drawImage(method='fft')`` it builds ``SB*`` profile objects
(e.g. ``Exponential._sbp = _galsim.SBExponential(r0, flux, gsparams._gsp)``), wraps numpy arrays in
``ImageView*`` objects by address, and calls, from ``GSObject.drawFFT`` and the ``_drawKImage``
recursion (galsim/gsobject.py:1927-2045, galsim/transform.py:592-600, galsim/convolve.py:420-426,
galsim/sum.py:341-347):

* ``SBProfile.drawK(ImageView, dk, jac_addr)`` on each leaf,
* ``ApplyKImagePhases(ImageView, imscale, jac_addr, cenx, ceny, flux_scaling)`` for shifts,
* ``wrapImage(ImageView, BoundsI, hermx, hermy)`` and ``irfft(in, out, shift_in, shift_out)``,
* ``goodFFTSize(n)``, and the profile scalars (``maxK``, ``stepK``, ``SersicHLR``, ...).

This module exports those names with the same argument order and meaning.  An ``SB*`` object
holds the equivalent galsim_amd profile; ``drawK`` lowers it (with the caller's Jacobian) to the
device program and evaluates it with the HIP kernels, ``ApplyKImagePhases``, ``wrapImage`` and
``irfft`` run HIP kernels, and every call copies between the caller's host buffer (the
``ImageView``'s address, step and stride) and the device.  Nothing falls back to the CPU: a
missing GPU or native library raises.  ``lower_to_gsmi`` gives the lowered program of any SB
tree for callers that drive the batched entry points themselves.

Limits: image views with positive step and stride; ``wrapImage`` for the drawFFT case
(hermx=True, hermy=False, wrapping (0..Nk/2, -Nk/2..Nk/2) onto (0..N/2, -N/2..N/2-1)); no photon
shooting.
"""
import ctypes

import numpy as np

from . import _lib
from . import render
from .gsparams import GSParams as _GSParams

__all__ = ['GSParams', 'BoundsI', 'PositionD', 'SBProfile', 'SBExponential', 'SBGaussian',
           'SBSersic', 'SBMoffat', 'SBKolmogorov', 'SBBox', 'SBTopHat', 'SBSpergel', 'SBAiry',
           'SBInclinedExponential', 'SBInclinedSersic', 'SBDeltaFunction', 'SBTransform',
           'SBConvolve', 'SBAutoConvolve', 'SBAutoCorrelate', 'SBAdd', 'SBDeconvolve',
           'ImageViewD', 'ImageViewF', 'ImageViewCD', 'ImageViewCF', 'ApplyKImagePhases',
           'wrapImage', 'irfft', 'goodFFTSize', 'SersicHLR', 'SersicIntegratedFlux',
           'SersicTruncatedScale', 'MoffatCalculateSRFromHLR', 'lower_to_gsmi',
           'Interpolant', 'Nearest', 'Linear', 'Cubic', 'Quintic', 'Lanczos', 'Delta',
           'SincInterpolant', 'SBInterpolatedImage', 'CalculateSizeContainingFlux']


# -- small value types (pysrc/Bounds.cpp, pysrc/SBProfile.cpp:54-57) ------------------------------
class GSParams:
    """GSParams(minimum_fft_size, maximum_fft_size, folding_threshold, stepk_minimum_hlr,
    maxk_threshold, kvalue_accuracy, xvalue_accuracy, table_spacing, realspace_relerr,
    realspace_abserr, integration_relerr, integration_abserr, shoot_accuracy)."""

    def __init__(self, *args):
        if len(args) != 13:
            raise TypeError("GSParams takes 13 arguments (%d given)" % len(args))
        self._p = _GSParams(*args)


def _gsp(g):
    return None if g is None else g._p


class BoundsI:
    def __init__(self, xmin, xmax, ymin, ymax):
        self.xmin, self.xmax, self.ymin, self.ymax = int(xmin), int(xmax), int(ymin), int(ymax)

    def _amd(self):
        from .position import BoundsI as B
        return B(self.xmin, self.xmax, self.ymin, self.ymax)


class PositionD:
    def __init__(self, x=0., y=0.):
        self.x, self.y = float(x), float(y)


# -- image views over caller memory (pysrc/Image.cpp:27-54) -------------------------------------
class _ImageView:
    _dtype = None

    def __init__(self, data_addr, step, stride, bounds):
        self._addr = int(data_addr)
        self._step = int(step)
        self._stride = int(stride)
        self._bounds = bounds
        if self._step <= 0 or self._stride <= 0:
            raise _lib.GalSimAMDError("image views need positive step and stride here")

    def _shape(self):
        b = self._bounds
        return b.ymax - b.ymin + 1, b.xmax - b.xmin + 1

    def _array(self):
        """numpy view of the caller's memory: element (x, y) at data + (x - xmin) step +
        (y - ymin) stride (include/galsim/Image.h:340-353)."""
        ny, nx = self._shape()
        isz = self._dtype.itemsize
        n = (nx - 1) * self._step + (ny - 1) * self._stride + 1
        buf = (ctypes.c_char * (n * isz)).from_address(self._addr)
        base = np.frombuffer(buf, dtype=self._dtype)
        return np.lib.stride_tricks.as_strided(base, shape=(ny, nx),
                                               strides=(self._stride * isz, self._step * isz))


class ImageViewD(_ImageView):
    _dtype = np.dtype(np.float64)


class ImageViewF(_ImageView):
    _dtype = np.dtype(np.float32)


class ImageViewCD(_ImageView):
    _dtype = np.dtype(np.complex128)


class ImageViewCF(_ImageView):
    _dtype = np.dtype(np.complex64)


def _jac(addr):
    """The 2x2 float64 Jacobian at a numpy array address (0: none)."""
    if not addr:
        return None
    buf = (ctypes.c_double * 4).from_address(int(addr))
    return np.array(buf, dtype=float).reshape(2, 2)


def _device():
    import torch
    if not torch.cuda.is_available():
        raise _lib.GalSimAMDError("galsim_amd draws on the GPU; no HIP device is available")
    return torch


# -- profiles (pysrc/SBProfile.cpp:25-73 and the per-profile modules) -------------------------
class SBProfile:
    """An SB profile: the galsim_amd object it stands for, in ``_obj``."""

    def __init__(self, obj):
        self._obj = obj

    def maxK(self):
        return self._obj.maxk

    def stepK(self):
        return self._obj.stepk

    def getFlux(self):
        return self._obj.flux

    def getPositiveFlux(self):
        return max(self._obj.flux, 0.)

    def getNegativeFlux(self):
        return max(-self._obj.flux, 0.)

    def kValue(self, pos):
        return self._obj.kValue(pos.x, pos.y)

    def drawK(self, image, dk, jac):
        """SBProfile::drawK (src/SBProfile.cpp:383-425): the k-image on the view's bounds, k =
        (x, y) dk mapped by the Jacobian at address ``jac`` (0: none), on the GPU."""
        if image._step != 1:
            raise _lib.GalSimAMDError("drawK needs a view with step 1 (as the reference)")
        b = image._bounds
        nk = None
        if b.xmin == 0 and b.ymin == -b.ymax and b.xmax == b.ymax:
            nk = 2 * b.xmax          # drawFFT_makeKImage's k-image: the exact clip rows apply
        arr = render.draw_kimage(self._obj, b._amd(), float(dk), _jac(jac), nk)
        image._array()[...] = arr.astype(image._dtype)

    def draw(self, image, dx, jac, xoff, yoff, flux_ratio):
        """SBProfile::draw (src/SBProfile.cpp:304-381): the profile at the centres of the
        view's pixels (coordinates x dx - xoff, mapped by the Jacobian at address ``jac``),
        times the pixel area and ``flux_ratio``, on the GPU (analytic leaves, their transforms
        and sums; real-space convolution raises)."""
        from . import real
        if image._step != 1:
            raise _lib.GalSimAMDError("draw needs a view with step 1 (as the reference)")
        arr = real.draw_real(self._obj, image._bounds._amd(), float(dx), jac=_jac(jac),
                             offset=(float(xoff), float(yoff)), flux_scaling=float(flux_ratio))
        image._array()[...] = arr.astype(image._dtype)

    def xValue(self, pos):
        """The surface brightness at ``pos`` (SBProfile::xValue), from a one-pixel draw."""
        from . import real
        from .position import BoundsI as B
        arr = real.draw_real(self._obj, B(0, 0, 0, 0), 1., offset=(-pos.x, -pos.y))
        return float(arr[0, 0])

    def shoot(self, *args):
        raise NotImplementedError("photon shooting is outside this library's path")


def _leaf(cls_name, factory):
    def init(self, *args):
        SBProfile.__init__(self, factory(*args))
    return type(cls_name, (SBProfile,), {'__init__': init})


def _amd():
    import galsim_amd
    return galsim_amd


SBExponential = _leaf('SBExponential', lambda r0, flux, gsp: _amd().Exponential(
    scale_radius=r0, flux=flux, gsparams=_gsp(gsp)))
SBGaussian = _leaf('SBGaussian', lambda sigma, flux, gsp: _amd().Gaussian(
    sigma=sigma, flux=flux, gsparams=_gsp(gsp)))
SBSersic = _leaf('SBSersic', lambda n, r0, flux, trunc, gsp: _amd().Sersic(
    n, scale_radius=r0, flux=flux, trunc=trunc, gsparams=_gsp(gsp)))
SBMoffat = _leaf('SBMoffat', lambda beta, r0, trunc, flux, gsp: _amd().Moffat(
    beta, scale_radius=r0, trunc=trunc, flux=flux, gsparams=_gsp(gsp)))
SBKolmogorov = _leaf('SBKolmogorov', lambda lor0, flux, gsp: _amd().Kolmogorov(
    lam_over_r0=lor0, flux=flux, gsparams=_gsp(gsp)))
SBBox = _leaf('SBBox', lambda w, h, flux, gsp: _amd().Box(w, h, flux=flux, gsparams=_gsp(gsp)))
SBTopHat = _leaf('SBTopHat', lambda r, flux, gsp: _amd().TopHat(r, flux=flux, gsparams=_gsp(gsp)))
SBSpergel = _leaf('SBSpergel', lambda nu, r0, flux, gsp: _amd().Spergel(
    nu, scale_radius=r0, flux=flux, gsparams=_gsp(gsp)))
SBAiry = _leaf('SBAiry', lambda lod, obs, flux, gsp: _amd().Airy(
    lam_over_diam=lod, obscuration=obs, flux=flux, gsparams=_gsp(gsp)))
SBInclinedExponential = _leaf(
    'SBInclinedExponential', lambda inc, r0, h0, flux, gsp: _amd().InclinedExponential(
        _amd().Angle(inc, _amd().radians), scale_radius=r0, scale_height=h0, flux=flux, gsparams=_gsp(gsp)))
SBInclinedSersic = _leaf(
    'SBInclinedSersic', lambda n, inc, r0, h0, flux, trunc, gsp: _amd().InclinedSersic(
        n, _amd().Angle(inc, _amd().radians), scale_radius=r0, scale_height=h0, flux=flux, trunc=trunc,
        gsparams=_gsp(gsp)))
SBDeltaFunction = _leaf('SBDeltaFunction', lambda flux, gsp: _amd().DeltaFunction(
    flux=flux, gsparams=_gsp(gsp)))


class SBTransform(SBProfile):
    """SBTransform(sbp, jac_addr, cenx, ceny, amp_scaling, gsp) (pysrc/SBTransform.cpp:25-31)."""

    def __init__(self, sbp, jac, cenx, ceny, amp_scaling, gsp):
        from .transform import Transformation
        SBProfile.__init__(self, Transformation(sbp._obj, _jac(jac), (cenx, ceny), amp_scaling,
                                                gsparams=_gsp(gsp), propagate_gsparams=False))


class SBConvolve(SBProfile):
    """SBConvolve(list, real_space, gsp).  Its k-image is the product of the factors' whatever
    real_space says (real_space only steers real-space drawing, which is not provided here)."""

    def __init__(self, slist, real_space, gsp):
        self._real_space = bool(real_space)
        SBProfile.__init__(self, _amd().Convolution([s._obj for s in slist], real_space=False,
                                                    gsparams=_gsp(gsp), propagate_gsparams=False))


class SBAutoConvolve(SBProfile):
    def __init__(self, sbp, real_space, gsp):
        SBProfile.__init__(self, _amd().AutoConvolution(sbp._obj, real_space=False,
                                                        gsparams=_gsp(gsp),
                                                        propagate_gsparams=False))


class SBAutoCorrelate(SBProfile):
    def __init__(self, sbp, real_space, gsp):
        SBProfile.__init__(self, _amd().AutoCorrelation(sbp._obj, real_space=False,
                                                        gsparams=_gsp(gsp),
                                                        propagate_gsparams=False))


class SBAdd(SBProfile):
    def __init__(self, slist, gsp):
        SBProfile.__init__(self, _amd().Sum([s._obj for s in slist], gsparams=_gsp(gsp),
                                            propagate_gsparams=False))


class SBDeconvolve(SBProfile):
    def __init__(self, sbp, gsp):
        SBProfile.__init__(self, _amd().Deconvolution(sbp._obj, gsparams=_gsp(gsp),
                                                      propagate_gsparams=False))


# -- interpolants and SBInterpolatedImage (pysrc/Interpolant.cpp:37-70,
#    pysrc/SBInterpolatedImage.cpp:25-37) ------------------------------------------------------
class Interpolant:
    """An interpolant: the galsim_amd object it stands for, in ``_i``."""

    def xval(self, x):
        return float(self._i.xval(x))

    def uval(self, u):
        return float(self._i.uval(u))

    def xvalMany(self, addr, n):
        """In place on n doubles at address ``addr``."""
        a = np.frombuffer((ctypes.c_double * int(n)).from_address(int(addr)), dtype=np.float64)
        a[:] = self._i.xval(a)

    def uvalMany(self, addr, n):
        a = np.frombuffer((ctypes.c_double * int(n)).from_address(int(addr)), dtype=np.float64)
        a[:] = self._i.uval(a)

    def urange(self):
        return self._i.urange


def _interp(cls_name, make):
    def init(self, *args):
        self._i = make(*args)
    return type(cls_name, (Interpolant,), {'__init__': init})


def _ia():
    from . import interpolant
    return interpolant


Nearest = _interp('Nearest', lambda gsp: _ia().Nearest(gsparams=_gsp(gsp)))
Linear = _interp('Linear', lambda gsp: _ia().Linear(gsparams=_gsp(gsp)))
Cubic = _interp('Cubic', lambda gsp: _ia().Cubic(gsparams=_gsp(gsp)))
Quintic = _interp('Quintic', lambda gsp: _ia().Quintic(gsparams=_gsp(gsp)))
Lanczos = _interp('Lanczos', lambda n, conserve_dc, gsp: _ia().Lanczos(
    n, conserve_dc, gsparams=_gsp(gsp)))


def _no_interp(name):
    def init(self, *args):
        raise NotImplementedError("the %s interpolant is not part of galsim_amd's "
                                  "InterpolatedImage" % name)
    return type(name, (Interpolant,), {'__init__': init})


Delta = _no_interp('Delta')
SincInterpolant = _no_interp('SincInterpolant')


class SBInterpolatedImage(SBProfile):
    """SBInterpolatedImage(image, init_bounds, nonzero_bounds, xInterp, kInterp, stepk, maxk,
    gsp): ``image`` a double view of the padded image on (-N/2..N/2-1)^2 (copied here), the
    bounds of the original image and of its nonzero part; stepk / maxk <= 0 are computed as the
    reference's constructor does.  Its k-image (the padded image's rfft) is taken on the GPU."""

    def __init__(self, image, init_bounds, nonzero_bounds, xInterp, kInterp, stepk, maxk, gsp):
        from .interpolatedimage import SBInterpolatedImageProfile
        b = image._bounds
        N = b.xmax - b.xmin + 1
        if b.xmin != -(N // 2) or b.ymin != b.xmin or b.ymax != b.xmax or N % 2:
            raise _lib.GalSimAMDError("SBInterpolatedImage needs the padded image on "
                                      "(-N/2..N/2-1)^2")
        ib, nb = init_bounds, nonzero_bounds
        SBProfile.__init__(self, SBInterpolatedImageProfile(
            np.array(image._array(), dtype=np.float64), (ib.xmin, ib.xmax, ib.ymin, ib.ymax),
            (nb.xmin, nb.xmax, nb.ymin, nb.ymax), xInterp._i, kInterp._i, float(stepk),
            float(maxk), _gsp(gsp)))

    def calculateMaxK(self, max_maxk=0.):
        self._obj.calculateMaxK(float(max_maxk))

    def calculateStepK(self, max_stepk=0.):
        self._obj.calculateStepK(float(max_stepk))


def CalculateSizeContainingFlux(image, target_flux):
    """CalculateSizeContainingFlux (src/SBInterpolatedImage.cpp:1084-1110) of a double view whose
    bounds contain (0, 0)."""
    from .interpolatedimage import _size_containing_flux
    b = image._bounds
    return _size_containing_flux(np.asarray(image._array(), dtype=np.float64), b.xmin, b.ymin,
                                 float(target_flux))


# -- free functions (pysrc/SBTransform.cpp:34-46, pysrc/Image.cpp:58-86) ----------------------
def _to_device(view, torch):
    a = np.ascontiguousarray(view._array(), dtype=np.complex128)
    return torch.from_numpy(a.view(np.float64)).cuda()


def ApplyKImagePhases(image, imscale, jac, cenx, ceny, flux_scaling):
    """ApplyKImagePhases (src/SBTransform.cpp:763-883) on the view, on the GPU."""
    torch = _device()
    d = _to_device(image, torch)
    ny, nx = image._shape()
    j = _jac(jac)
    jp = None if j is None else np.ascontiguousarray(j.reshape(4))
    _lib.check(_lib.lib().gsmi_apply_kimage_phases_f64(
        d.data_ptr(), image._bounds.xmin, image._bounds.ymin, nx, ny, float(imscale),
        None if jp is None else jp.ctypes.data, float(cenx), float(ceny), float(flux_scaling),
        torch.cuda.current_stream().cuda_stream), "ApplyKImagePhases")
    h = d.cpu().numpy().view(np.complex128).reshape(ny, nx)
    image._array()[...] = h.astype(image._dtype)


def wrapImage(image, bounds, hermx, hermy):
    """wrapImage (src/Image.cpp:577-726) in place, for drawFFT_finish's call: hermx, the view's
    bounds (0..Nk/2, -Nk/2..Nk/2) folded onto ``bounds`` = (0..N/2, -N/2..N/2-1)."""
    b = image._bounds
    nk = 2 * b.xmax
    N = 2 * bounds.xmax
    if not (hermx and not hermy and b.xmin == 0 and b.ymin == -nk // 2 and b.ymax == nk // 2
            and bounds.xmin == 0 and bounds.ymin == -N // 2 and bounds.ymax == N // 2 - 1
            and N <= nk):
        raise NotImplementedError("wrapImage here covers drawFFT's Hermitian-x wrap only")
    torch = _device()
    d = _to_device(image, torch)
    out = torch.empty((N, N // 2 + 1, 2), dtype=torch.float64, device='cuda')
    _lib.check(_lib.lib().gsmi_wrap_hermx_f64(d.data_ptr(), nk, N, out.data_ptr(),
                                              torch.cuda.current_stream().cuda_stream),
               "wrapImage")
    h = out.cpu().numpy().view(np.complex128).reshape(N, N // 2 + 1)
    # rows ky = -N/2..N/2-1 sit at view rows (ky + Nk/2), columns 0..N/2
    r0 = nk // 2 - N // 2
    image._array()[r0:r0 + N, :N // 2 + 1] = h.astype(image._dtype)


def irfft(kin, out, shift_in, shift_out):
    """irfft (src/Image.cpp:814-911): kin (0..N/2, -N/2..N/2-1) complex -> out (-N/2..N/2+1,
    -N/2..N/2-1) float64 (the two columns past N/2-1 are FFTW's padding; written as zero)."""
    b = kin._bounds
    N = 2 * b.xmax
    ob = out._bounds
    if not (b.xmin == 0 and b.ymin == -N // 2 and b.ymax == N // 2 - 1 and
            ob.xmin == -N // 2 and ob.xmax == N // 2 + 1 and ob.ymin == -N // 2 and
            ob.ymax == N // 2 - 1):
        raise _lib.GalSimAMDError("irfft: bounds must be (0..N/2, -N/2..N/2-1) -> "
                                  "(-N/2..N/2+1, -N/2..N/2-1)")
    if out._dtype != np.float64:
        raise _lib.GalSimAMDError("irfft writes a float64 image")
    torch = _device()
    from .fft import irfft_scratch_rows
    d = _to_device(kin, torch)
    real = torch.empty((N, N), dtype=torch.float64, device='cuda')
    tmp = torch.empty((irfft_scratch_rows(N), N // 2 + 1), dtype=torch.complex128,
                      device='cuda')
    _lib.check(_lib.lib().gsmi_irfft2_f64(d.data_ptr(), N, N, int(bool(shift_in)),
                                          int(bool(shift_out)), real.data_ptr(), tmp.data_ptr(),
                                          torch.cuda.current_stream().cuda_stream), "irfft")
    o = out._array()
    o[:, :N] = real.cpu().numpy()
    o[:, N:] = 0.


def goodFFTSize(n):
    return int(_lib.lib().gsmi_good_fft_size(int(n)))


def SersicHLR(n, flux_fraction):
    out = np.zeros(1)
    _lib.check(_lib.lib().gsmi_sersic_hlr(float(n), float(flux_fraction), _lib.ptr(out)),
               "SersicHLR")
    return float(out[0])


def SersicIntegratedFlux(n, r):
    out = np.zeros(1)
    _lib.check(_lib.lib().gsmi_sersic_integrated_flux(float(n), float(r), _lib.ptr(out)),
               "SersicIntegratedFlux")
    return float(out[0])


def SersicTruncatedScale(n, hlr, trunc):
    out = np.zeros(1)
    _lib.check(_lib.lib().gsmi_sersic_truncated_scale(float(n), float(hlr), float(trunc),
                                                      _lib.ptr(out)), "SersicTruncatedScale")
    return float(out[0])


def MoffatCalculateSRFromHLR(hlr, trunc, beta):
    out = np.zeros(1)
    if trunc == 0.:
        rc = _lib.lib().gsmi_moffat_scale_radius_from_hlr(float(hlr), float(beta), _lib.ptr(out))
    else:
        rc = _lib.lib().gsmi_moffat_scale_radius_from_hlr_trunc(float(hlr), float(trunc),
                                                                float(beta), _lib.ptr(out))
    _lib.check(rc, "MoffatCalculateSRFromHLR")
    return float(out[0])


def lower_to_gsmi(sbp, N, Nk, dk, precision='f64'):
    """The lowered k-space program of an SB tree for the batched entry points of
    include/gsmi.h: (ops, leaves, tables, tables32, cliprows) as numpy record arrays for one
    stamp of FFT size N (k-image Nk, spacing dk in pixel units), with the separable factors
    hoisted behind the program (op_count, axis_count also returned)."""
    ctx = render.Lowering(precision, Nk=Nk, dk=dk)
    sbp._obj._lower(ctx, None)
    nops, naxis = render.hoist_axis_factors(ctx, 0)
    render.check_stack(ctx.ops, 0, nops, precision)
    tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
    clip = np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, dtype=np.int32)
    return dict(ops=ctx.ops_array(), leaves=ctx.leaves_array(), tables=tables,
                tables32=ctx.tables32_array(), cliprows=clip, op_count=nops, axis_count=naxis,
                N=N, Nk=Nk, dk=dk)
