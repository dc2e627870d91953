"""InterpolatedImage (galsim/interpolatedimage.py; src/SBInterpolatedImage.cpp) on the FFT path.

Construction follows the reference: the image is re-centred on (0, 0), zero-padded to
goodFFTSize(pad_factor * size), its flux, stepk (the box holding 1 - folding_threshold of the
flux, in quadrature with the x-interpolant's range) and maxk (the last k row of the padded
image's transform above maxk_threshold of the flux) are computed, and the drawn profile is
SBTransform(SBInterpolatedImage, pixel-scale Jacobian, -offset, flux ratio).  The padded image's
rfft -- the k-image the interpolation reads -- is taken on the GPU once per object.

On the device the profile is one GSMI_LEAF_INTERP_IMAGE leaf: each k sample sums the
k-interpolant's taps over that k-image and multiplies by the x-interpolant's transform, then the
transform's phase (or flux scaling) follows as program ops, as SBTransform::fillKImage applies
them after its adaptee's fill."""
import math

import numpy as np

from . import _lib
from .gsobject import GSObject
from .gsparams import GSParams
from .image import Image
from .interpolant import Quintic, convert_interpolant
from .position import PositionD


class SBInterpolatedImageProfile(GSObject):
    """SBInterpolatedImage itself (src/SBInterpolatedImage.cpp:95-152), in the image's pixel
    coordinates with no transformation: the padded image ``xim`` (array on (-N/2..N/2-1)^2),
    the original image's bounds ``init_bounds`` and the nonzero region ``nonzero_bounds`` (each
    (xmin, xmax, ymin, ymax)), the two interpolants, and stepk / maxk (<= 0: computed as the
    constructor does -- the image box in quadrature with the x-interpolant's range, and the
    x-interpolant's own maxk).  ``kimage`` (optional) is the padded image's rfft, if already
    taken."""
    _is_analytic_x = True

    def __init__(self, xim, init_bounds, nonzero_bounds, x_interpolant, k_interpolant, stepk,
                 maxk, gsparams=None, kimage=None):
        self._gsparams = GSParams.check(gsparams)
        self._xim = np.ascontiguousarray(xim, dtype=np.float64)
        if self._xim.ndim != 2 or self._xim.shape[0] != self._xim.shape[1] or \
                self._xim.shape[0] % 2:
            raise ValueError("the padded image must be square with an even side")
        self._init_bounds = tuple(int(v) for v in init_bounds)
        self._nonzero_bounds = tuple(int(v) for v in nonzero_bounds)
        self._x_interpolant = x_interpolant
        self._k_interpolant = k_interpolant
        self._maxk1 = x_interpolant.urange * 2. * math.pi
        if stepk <= 0.:
            xmin, xmax, ymin, ymax = self._init_bounds
            R = max((xmax - xmin) / 2., (ymax - ymin) / 2.)
            R2 = x_interpolant.xrange
            stepk = math.pi / math.sqrt(R * R + R2 * R2)
        self._stepk = float(stepk)
        self._maxk = float(maxk) if maxk > 0. else self._maxk1
        self._kimage = kimage

    def _sub(self, b):
        h = self._xim.shape[0] // 2
        xmin, xmax, ymin, ymax = b
        return self._xim[ymin + h: ymax + h + 1, xmin + h: xmax + h + 1]

    @property
    def _flux(self):
        """getFlux (:1064-1078): the sum over the nonzero image."""
        return float(np.sum(self._sub(self._nonzero_bounds), dtype=np.float64))

    @property
    def kimage(self):
        """rfft(shift_in, shift_out) of the padded image on (0..N/2, -N/2..N/2-1), array
        [ky + N/2, kx] (checkK, :450-464), taken on the GPU once."""
        if self._kimage is None:
            from .fft import rfft2
            self._kimage = np.ascontiguousarray(rfft2(self._xim, shift_in=True, shift_out=True),
                                                dtype=np.complex128)
        return self._kimage

    def calculateMaxK(self, max_maxk=0.):
        """calculateMaxK (:1146-1217); sets and returns maxk."""
        if max_maxk == 0.:
            max_maxk = self._maxk
        self._maxk = _calculate_maxk(self.kimage, self._flux, self._gsparams.maxk_threshold,
                                     max_maxk)
        return self._maxk

    def calculateStepK(self, max_stepk=0.):
        """calculateStepK (:1117-1144): the box holding 1 - folding_threshold of the flux, in
        quadrature with the x-interpolant's range; sets and returns stepk."""
        b = self._init_bounds
        thresh = (1. - self._gsparams.folding_threshold) * self._flux
        R = _size_containing_flux(self._sub(b), b[0], b[2], thresh)
        R2 = self._x_interpolant.xrange
        self._stepk = math.pi / math.sqrt(R * R + R2 * R2)
        return self._stepk

    def _lower(self, ctx, jac):
        self._lower_leaf(ctx, jac, None)

    def _lower_leaf(self, ctx, jac, J):
        """The leaf as SBInterpolatedImage::fillKImage gets it from SBProfile::drawK with the
        Jacobian ``jac``, through an SBTransform of diagonal Jacobian ``J`` if given
        (src/SBProfile.cpp:383-425, src/SBTransform.cpp:885-969): the separable fill with its
        column / row range while every Jacobian is diagonal, else the fill with cross terms; in
        a kValue lowering, kValue's form."""
        from .render import LeafRecord
        A = D = 1.
        if J is not None:
            A, D = J[0, 0], J[1, 1]
        jac1 = J if jac is None else jac if J is None else jac.dot(J)
        j1 = np.eye(2) if jac1 is None else jac1
        sep = jac is None or (jac[0, 1] == 0. and jac[1, 0] == 0.)
        kv = bool(getattr(ctx, 'kvalue_mode', False))
        rec = LeafRecord()
        rec['type'] = _lib.LEAF_INTERP_IMAGE
        rec['separable'] = 1 if sep else 0
        rec['kind'] = 1 if kv else 0
        rec['jac'] = (j1[0, 0], j1[0, 1], j1[1, 0], j1[1, 1])
        rec['flux'] = 1.
        K = self.kimage
        No2 = K.shape[1] - 1
        kscale = No2 / math.pi
        xi, ki = self._x_interpolant, self._k_interpolant
        maxk1 = self._maxk1
        blk = np.zeros(_lib.INTERP_BLOCK)
        blk[_lib.INTERP_N2] = No2
        blk[_lib.INTERP_KSCALE] = kscale
        blk[_lib.INTERP_MAXK1] = maxk1
        blk[_lib.INTERP_MAXK1S] = maxk1 * kscale
        blk[_lib.INTERP_KXR] = ki.xrange
        if sep and not kv:
            xmin, ymin, m, n = _kgrid(ctx)
            dk = float(ctx.dk)
            # SBProfile::drawK (src/SBProfile.cpp:392-410), then SBTransform's fwdT
            x0, y0 = xmin * dk, ymin * dk
            if jac is None:
                kx0, dkx, ky0, dky = x0, dk, y0, dk
            else:
                kx0, dkx, ky0, dky = x0 * jac[0, 0], dk * jac[0, 0], y0 * jac[1, 1], \
                    dk * jac[1, 1]
            kx0, dkx, ky0, dky = A * kx0, A * dkx, D * ky0, D * dky
            i1, i2, j1_, j2 = _fill_range(maxk1, kx0, dkx, ky0, dky, m, n)
            blk[_lib.INTERP_I1], blk[_lib.INTERP_I2] = i1, i2
            blk[_lib.INTERP_J1], blk[_lib.INTERP_J2] = j1_, j2
            blk[_lib.INTERP_XMIN], blk[_lib.INTERP_YMIN] = xmin, ymin
        xcode, xn, _, _, xumax, xtab = xi._device()
        blk[_lib.INTERP_XCODE] = xcode
        blk[_lib.INTERP_XN] = xn
        blk[_lib.INTERP_XUMAX] = xumax
        blk[_lib.INTERP_XUTAB] = -1
        if xtab is not None:
            blk[_lib.INTERP_XUTAB] = ctx.add_block(('lanczos_u', xi._key()), xtab)
            blk[_lib.INTERP_XUTABN] = len(xtab) // 3
        kcode, kn, kdc, kK, _, _ = ki._device()
        blk[_lib.INTERP_KCODE] = kcode
        blk[_lib.INTERP_KN] = kn
        blk[_lib.INTERP_KDC] = 1. if kdc else 0.
        blk[_lib.INTERP_KK: _lib.INTERP_KK + 5] = kK
        blk[_lib.INTERP_KIMAGE] = ctx.add_block(('interp_kimage', id(K)), K.view(np.float64),
                                                owner=K)
        rec['table'] = ctx.add_block(None, blk)
        rec['ntab'] = No2
        ctx.emit(_lib.OP_LEAF, ctx.add_leaf(rec))


class InterpolatedImage(GSObject):
    """galsim.InterpolatedImage(image, x_interpolant=None, k_interpolant=None,
    normalization='flux', scale=None, flux=None, pad_factor=4., calculate_stepk=True,
    calculate_maxk=True, use_true_center=True, offset=None, gsparams=None).

    ``image`` is a galsim_amd Image (or a 2-d array, with ``scale``).  The pixel scale is the wcs
    (a general wcs, noise padding, a pad image and depixelize are not part of galsim_amd)."""
    _is_analytic_x = True
    _has_hard_edges = False

    def __init__(self, image, x_interpolant=None, k_interpolant=None, normalization='flux',
                 scale=None, wcs=None, flux=None, pad_factor=4., noise_pad_size=0, noise_pad=0.,
                 rng=None, pad_image=None, calculate_stepk=True, calculate_maxk=True,
                 use_cache=True, use_true_center=True, depixelize=False, offset=None,
                 gsparams=None, _force_stepk=0., _force_maxk=0., hdu=None):
        if isinstance(image, np.ndarray):
            image = Image(image)
        if not isinstance(image, Image):
            raise TypeError("Supplied image must be an Image")
        if not image.bounds.isDefined():
            raise _lib.GalSimAMDUndefinedBoundsError("Supplied image does not have bounds "
                                                     "defined.")
        if normalization.lower() not in ("flux", "f", "surface brightness", "sb"):
            raise ValueError("Invalid normalization requested: %r" % normalization)
        if wcs is not None or noise_pad_size or noise_pad or pad_image is not None or \
                depixelize or hdu is not None:
            raise NotImplementedError("InterpolatedImage with a general wcs, noise padding, a "
                                      "pad image or depixelize is not part of galsim_amd")
        self._gsparams = GSParams.check(gsparams)
        self._x_interpolant = (Quintic(gsparams=self._gsparams) if x_interpolant is None else
                               convert_interpolant(x_interpolant).withGSParams(self._gsparams))
        self._k_interpolant = (Quintic(gsparams=self._gsparams) if k_interpolant is None else
                               convert_interpolant(k_interpolant).withGSParams(self._gsparams))
        if scale is None:
            scale = image.scale
        if scale is None:
            raise ValueError("No information given with Image or keywords about pixel scale!")
        self._scale = float(scale)
        arr = np.ascontiguousarray(np.asarray(image.array).real, dtype=np.float64)
        ny, nx = arr.shape
        # setCenter(0, 0): the image's centre pixel moves to the origin
        b = image.bounds
        cx, cy = b.center.x, b.center.y
        self._bx0, self._by0 = b.xmin - cx, b.ymin - cy
        self._nx, self._ny = nx, ny
        # _adjust_offset (galsim/gsobject.py:1234-1253) of the original image's bounds
        off = PositionD(0., 0.) if offset is None else PositionD(offset)
        off = GSObject._adjust_offset(b, off, None, use_true_center)
        self._offset = (off.x, off.y)
        # _buildRealImage (galsim/interpolatedimage.py:447-519): zero padding to goodFFTSize
        if pad_factor <= 0.:
            raise ValueError("Invalid pad_factor <= 0 in InterpolatedImage")
        pad_size = max(ny, nx)
        if pad_factor > 1.:
            pad_size = int(math.ceil(pad_factor * pad_size))
        pad_size = Image.good_fft_size(pad_size)
        self._pad_factor = pad_factor
        h = pad_size // 2
        xim = np.zeros((pad_size, pad_size))
        xim[self._by0 + h: self._by0 + h + ny, self._bx0 + h: self._bx0 + h + nx] = arr
        self._xim = xim
        self._array = arr
        self._image_flux = float(np.sum(arr, dtype=np.float64))
        if (calculate_stepk or calculate_maxk or flux is not None) and self._image_flux == 0.:
            raise ValueError("This input image has zero total flux. It does not define a valid "
                             "surface brightness profile.")
        # _getFlux (:568-575)
        if flux is None:
            flux = self._image_flux
            if normalization.lower() in ('surface brightness', 'sb'):
                flux *= self._scale * self._scale
        self._flux = float(flux)
        self._kimage = None
        self._calculate_stepk = calculate_stepk
        self._calculate_maxk = calculate_maxk
        self._stepk = self._get_stepk(calculate_stepk, _force_stepk)
        self._maxk = self._get_maxk(calculate_maxk, _force_maxk)

    # -- set-up (galsim/interpolatedimage.py:577-628) ------------------------------------------
    def _get_stepk(self, calculate_stepk, force):
        if force > 0.:
            return force
        if calculate_stepk:
            thresh = (1. - self._gsparams.folding_threshold) * self._image_flux
            R = _size_containing_flux(self._array, self._bx0, self._by0, thresh)
        else:
            R = max(self._array.shape) / 2. - 0.5
        R = math.hypot(R, self._x_interpolant.xrange)
        return math.pi / (R * self._scale)

    def _get_maxk(self, calculate_maxk, force):
        if force > 0.:
            return force
        if calculate_maxk:
            # SBInterpolatedImage's own maxk (the x-interpolant's urange) caps the search
            maxk1 = self._x_interpolant.urange * 2. * math.pi
            max_maxk = maxk1 if calculate_maxk is True else float(calculate_maxk)
            return _calculate_maxk(self.kimage, self._image_flux,
                                   self._gsparams.maxk_threshold, max_maxk) / self._scale
        return self._x_interpolant.krange / self._scale

    @property
    def kimage(self):
        """The padded image's rfft(shift_in, shift_out) on (0..N/2, -N/2..N/2-1), array
        [ky + N/2, kx] (SBInterpolatedImage::checkK, src/SBInterpolatedImage.cpp:450-464),
        computed on the GPU."""
        if self._kimage is None:
            from .fft import rfft2
            self._kimage = np.ascontiguousarray(rfft2(self._xim, shift_in=True, shift_out=True),
                                                dtype=np.complex128)
        return self._kimage

    # -- GSObject properties -------------------------------------------------------------------
    @property
    def x_interpolant(self):
        return self._x_interpolant

    @property
    def k_interpolant(self):
        return self._k_interpolant

    @property
    def image(self):
        im = Image(self._array.copy(), xmin=self._bx0, ymin=self._by0)
        im.scale = self._scale
        return im

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        ret._x_interpolant = self._x_interpolant.withGSParams(ret._gsparams)
        ret._k_interpolant = self._k_interpolant.withGSParams(ret._gsparams)
        if ret._gsparams.folding_threshold != self._gsparams.folding_threshold:
            ret._stepk = ret._get_stepk(self._calculate_stepk, 0.)
        if ret._gsparams.maxk_threshold != self._gsparams.maxk_threshold:
            ret._maxk = ret._get_maxk(self._calculate_maxk, 0.)
        return ret

    def _world(self):
        """(J, (cx, cy), flux scaling) of the SBTransform the reference draws: the shift by
        -offset composed with the pixel-scale Jacobian (galsim/interpolatedimage.py:398-427,
        galsim/wcs.py:1538-1548, galsim/transform.py:604-640)."""
        s = self._scale
        ratio = self._flux / self._image_flux if self._flux != self._image_flux else 1.
        inv = 1. / s
        if s == 1.:
            J = None
            cx, cy = -self._offset[0], -self._offset[1]
            det = 1.
        else:
            J = np.array([[s, 0.], [0., s]])
            # _fwd_normal of the shift through J
            cx, cy = -self._offset[0] * s, -self._offset[1] * s
            det = s * s
        return J, (cx, cy), abs(det) * (inv * inv * ratio)

    # -- lowering ------------------------------------------------------------------------------
    def _sb(self):
        """The SBInterpolatedImage it draws through (stepk and maxk in pixel units)."""
        b = (self._bx0, self._bx0 + self._nx - 1, self._by0, self._by0 + self._ny - 1)
        return SBInterpolatedImageProfile(self._xim, b, b, self._x_interpolant,
                                          self._k_interpolant, self._stepk * self._scale,
                                          self._maxk * self._scale, self._gsparams,
                                          kimage=self.kimage)

    def _lower(self, ctx, jac):
        """InterpolatedImage._drawKImage (galsim/interpolatedimage.py:733-735): _sbp.drawK ->
        SBTransform::fillKImage (src/SBTransform.cpp:885-969): the SBInterpolatedImage's fill
        with the k grid the transform hands it, then the shift's phases or the flux scaling."""
        J, (cx, cy), fs = self._world()
        self._sb()._lower_leaf(ctx, jac, J)
        if cx != 0. or cy != 0.:
            if jac is None:
                ux, uy = cx, cy
            else:
                ux = jac[0, 0] * cx + jac[0, 1] * cy
                uy = jac[1, 0] * cx + jac[1, 1] * cy
            ctx.emit(_lib.OP_PHASE, 0, ux, uy, fs)
        elif abs(fs - 1.) > self._gsparams.kvalue_accuracy:
            ctx.emit(_lib.OP_SCALE, 0, fs)

    def __repr__(self):
        return ('galsim_amd.InterpolatedImage(%dx%d image, scale=%r, flux=%r, x_interpolant=%r, '
                'k_interpolant=%r)' % (self._nx, self._ny, self._scale, self._flux,
                                       self._x_interpolant, self._k_interpolant))


def _kgrid(ctx):
    """(xmin, ymin, columns, rows) of the k-image being filled: the lowering's own, else
    drawFFT_makeKImage's (0..Nk/2, -Nk/2..Nk/2)."""
    kb = getattr(ctx, 'kbounds', None)
    if kb is not None:
        return kb
    if ctx.Nk is None:
        raise _lib.GalSimAMDError("InterpolatedImage lowering needs the k-image's bounds")
    Nk = int(ctx.Nk)
    return 0, -(Nk // 2), Nk // 2 + 1, Nk + 1


def _fill_range(maxk1, kx0, dkx, ky0, dky, m, n):
    """The columns [i1, i2) and rows [j1, j2) SBInterpolatedImage's separable fillKImage
    evaluates (src/SBInterpolatedImage.cpp:783-797); an empty range as (0, 0, 0, 0)."""
    absdkx, absdky = abs(dkx), abs(dky)
    i1 = max(int(-maxk1 / absdkx - kx0 / dkx), 0)
    i2 = min(int(maxk1 / absdkx - kx0 / dkx) + 1, m)
    j1 = max(int(-maxk1 / absdky - ky0 / dky), 0)
    j2 = min(int(maxk1 / absdky - ky0 / dky) + 1, n)
    if i1 >= m or i2 < 0 or j1 >= n or j2 < 0 or i1 >= i2 or j1 >= j2:
        return 0, 0, 0, 0
    return i1, i2, j1, j2


def _size_containing_flux(arr, xmin, ymin, target):
    """CalculateSizeContainingFlux (src/SBInterpolatedImage.cpp:1084-1110): the half-size + 0.5
    of the smallest centred box whose flux reaches ``target``, growing one ring at a time."""
    ny, nx = arr.shape
    xmax, ymax = xmin + nx - 1, ymin + ny - 1
    p = 1. if target > 0. else -1.
    dmax = min((xmax - xmin) // 2, (ymax - ymin) // 2)
    flux = arr[-ymin, -xmin]
    d = 1
    while d <= dmax:
        ring = arr[-d - ymin: d - ymin + 1, -d - xmin: d - xmin + 1]
        flux += ring[0, :].sum() + ring[-1, :].sum() + ring[1:-1, 0].sum() + \
            ring[1:-1, -1].sum()
        if p * flux >= p * target:
            break
        d += 1
    return d + 0.5


def _calculate_maxk(K, flux, maxk_threshold, max_maxk):
    """SBInterpolatedImage::calculateMaxK (src/SBInterpolatedImage.cpp:1146-1217): scanning
    squares of half-size ix = 0, 1, ... over the stored half-plane, the last one with a
    |k-value|^2 above (maxk_threshold flux)^2; stops after 5 squares below it."""
    No2 = K.shape[1] - 1
    dk = math.pi / No2
    thresh = maxk_threshold * flux
    thresh *= thresh
    nrm = K.real * K.real + K.imag * K.imag          # [q + No2, p]
    maxk_ix = 0
    n_below = 0
    max_ix = min(int(math.ceil(max_maxk / dk)), No2)
    for ix in range(max_ix + 1):
        iy = np.arange(ix + 1)
        above = nrm[No2 - ix, iy] > thresh
        if ix != No2:
            # the upper side, read where the lower one is below threshold (iy != ix)
            above |= (iy != ix) & (nrm[(ix + No2) % (2 * No2), iy] > thresh)
        above |= (iy > 0) & (nrm[No2 - iy, ix] > thresh)
        if ix > 0:
            up = np.where(iy != No2, (iy + No2) % (2 * No2), 0)
            above |= (iy != No2) & (nrm[up, ix] > thresh)
        if above.any():
            maxk_ix = ix
            n_below = 0
        n_below += 1
        if n_below == 5:
            break
    return (maxk_ix + 1) * dk
