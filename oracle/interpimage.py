"""InterpolatedImage on the FFT path, restated in numpy FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Restates:
* the interpolants' xval / uval / ranges (src/Interpolant.cpp:167-560, 563-919;
  include/galsim/Interpolant.h:224-521; galsim/interpolant.py) -- Nearest, Linear, Cubic,
  Quintic and Lanczos(n, conserve_dc);
* SBInterpolatedImage: construction (src/SBInterpolatedImage.cpp:95-152), the k-image
  (checkK, :450-464: rfft with shift_in and shift_out), kValue (:385-448), both fillKImage forms
  (:767-936 with izero/jzero, :938-1009 with cross terms), the wrapped k-index walk of
  KValueInnerLoop (:217-222, :317-383), calculateMaxK (:1146-1217), getFlux (:1064-1078) and
  CalculateSizeContainingFlux (:1084-1110);
* galsim.InterpolatedImage's set-up (galsim/interpolatedimage.py:303-428, 447-519, 568-628):
  re-centring, offset, zero padding to goodFFTSize, flux normalisation, stepk / maxk, and the
  profile _sbp = SBTransform(SBInterpolatedImage, wcs Jacobian, -offset, flux ratio);
* the SBTransform / SBProfile::drawK dispatch that hands the leaf its k grid
  (src/SBProfile.cpp:383-425, src/SBTransform.cpp:885-969).

The k-interpolation sums are evaluated as numpy gathers over all taps at once; the reference
sums the taps in sequence (and, for p < 0, with a conjugating dot product), so values agree to
rounding, not bit for bit.
"""
import math

import numpy as np

from . import draw as dr
from . import numerics as nm
from .fft import rfft2

SMALL = 10. * np.finfo(float).eps
USCALE = 1. / (2. * math.pi)


def sinc(x):
    """math::sinc (src/math/Sinc.cpp:29-33): sin(pi x) / (pi x), 1 - pi^2 x^2 / 6 below 1e-4."""
    x = np.asarray(x, dtype=float)
    with np.errstate(invalid='ignore', divide='ignore'):
        big = np.sin(np.pi * x) / (np.pi * x)
    return np.where(np.abs(x) < 1.e-4, 1. - (np.pi * np.pi / 6.) * x * x, big)


def si(x):
    """Sine integral Si(x) (src/math/Sinc.cpp:101-170 computes it by Pade approximants good to
    1e-16; here scipy's sici, which is accurate to the same level)."""
    from scipy.special import sici
    return sici(x)[0]


# ----------------------------------------------------------------------------------------------
# Interpolants
# ----------------------------------------------------------------------------------------------
class Interpolant:
    name = None

    def __init__(self, gsparams=None):
        from . import profiles as pr
        self.gsparams = gsparams or pr.GSParams()

    @property
    def krange(self):
        """galsim/interpolant.py: 2 pi urange."""
        return 2. * math.pi * self.urange


class Nearest(Interpolant):
    """src/Interpolant.cpp:171-176; include/galsim/Interpolant.h:276-278."""
    name = 'nearest'
    xrange = 0.5
    ixrange = 1

    @property
    def urange(self):
        return 1. / (math.pi * self.gsparams.kvalue_accuracy)

    def xval(self, x):
        return np.where(np.abs(x) > 0.5, 0., 1.)

    def uval(self, u):
        return sinc(u)


class Linear(Interpolant):
    """src/Interpolant.cpp:245-255; include/galsim/Interpolant.h:345-347."""
    name = 'linear'
    xrange = 1.
    ixrange = 2

    @property
    def urange(self):
        return math.sqrt(1. / self.gsparams.kvalue_accuracy) / math.pi

    def xval(self, x):
        x = np.abs(x)
        return np.where(x > 1., 0., 1. - x)

    def uval(self, u):
        s = sinc(u)
        return s * s


class Cubic(Interpolant):
    """src/Interpolant.cpp:284-370 (USE_TABLES is off: closed-form uval and uMax)."""
    name = 'cubic'
    xrange = 2.
    ixrange = 4

    @property
    def urange(self):
        return (3. * math.sqrt(3.) / 8. / self.gsparams.kvalue_accuracy) ** (1. / 3.) / math.pi

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        inner = 1. + x * x * (1.5 * x - 2.5)
        outer = -0.5 * (x - 1.) * (x - 2.) * (x - 2.)
        return np.where(x < 1., inner, np.where(x < 2., outer, 0.))

    def uval(self, u):
        u = np.abs(u)
        s = sinc(u)
        c = np.cos(np.pi * u)
        return s * s * s * (3. * s - 2. * c)


class Quintic(Interpolant):
    """src/Interpolant.cpp:389-531 (the non-ALT_QUINTIC form; USE_TABLES is off)."""
    name = 'quintic'
    xrange = 3.
    ixrange = 6

    @property
    def urange(self):
        return (25. * math.sqrt(5.) / 108. / self.gsparams.kvalue_accuracy) ** (1. / 3.) / math.pi

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        r1 = 1. + x * x * x * (-95. / 12. + x * (23. / 2. + x * (-55. / 12.)))
        r2 = (x - 1.) * (x - 2.) * (-23. / 4. + x * (29. / 2. + x * (-83. / 8. + x * (55. / 24.))))
        r3 = (x - 2.) * (x - 3.) * (x - 3.) * (-9. / 4. + x * (25. / 12. + x * (-11. / 24.)))
        return np.where(x <= 1., r1, np.where(x <= 2., r2, np.where(x <= 3., r3, 0.)))

    def uval(self, u):
        u = np.abs(u)
        s = sinc(u)
        piu = np.pi * u
        c = np.cos(piu)
        ssq = s * s
        piusq = piu * piu
        return s * ssq * ssq * (s * (55. - 19. * piusq) + 2. * c * (piusq - 27.))


class Lanczos(Interpolant):
    """src/Interpolant.cpp:563-919: xval from xCalc (sin(pi x) sin(pi x / n) n / (pi x)^2 with the
    conserve_dc correction), uval from the spline table of uCalc built on u = 0, du, ...
    (the reference's own table, Table::spline)."""
    name = 'lanczos'

    def __init__(self, n, conserve_dc=True, gsparams=None):
        super().__init__(gsparams)
        self.n = int(n)
        self.nd = float(n)
        self.conserve_dc = bool(conserve_dc)
        self.xrange = self.nd
        self.ixrange = 2 * self.n
        K = [0.] + [self._ucalc_raw(float(j)) for j in range(1, 6)]
        self.K = K
        self.C = [1. + 2. * (K[1] * (1. + 3. * K[1] + K[2] + K[3]) + K[2] + K[3] + K[4] + K[5]),
                  -K[1] * (1. + 4. * K[1] + K[2] + 2. * K[3]),
                  K[1] * (K[1] - 2. * K[2] + K[3]) - K[2],
                  K[1] * (K[1] - 2. * K[3]) - K[3],
                  K[1] * K[3] - K[4],
                  -K[5]]
        tol = self.gsparams.kvalue_accuracy
        ustep = self.gsparams.table_spacing * (tol / 200.) ** 0.25 / self.nd
        args, vals = [], []
        umax = 0.
        u = 0.
        while u - umax < 1. / self.nd or u < 1.1:
            v = self._ucalc(u)
            args.append(u)
            vals.append(v)
            if abs(v) > tol:
                umax = u
            u += ustep
        self.utab = nm.SplineTable(args, vals)
        self.umax = umax

    @property
    def urange(self):
        return self.umax

    def _ucalc_raw(self, u):
        vp = self.nd * (2. * u + 1.)
        vm = self.nd * (2. * u - 1.)
        r = ((vm - 1.) * si(math.pi * (vm - 1.)) - (vm + 1.) * si(math.pi * (vm + 1.))
             - (vp - 1.) * si(math.pi * (vp - 1.)) + (vp + 1.) * si(math.pi * (vp + 1.)))
        return r / (2. * math.pi)

    def _ucalc(self, u):
        r = self._ucalc_raw(u)
        if self.conserve_dc:
            C = self.C
            r *= C[0]
            for j in range(1, 6):
                r += C[j] * (self._ucalc_raw(u + j) + self._ucalc_raw(u - j))
        return r

    def xval(self, x):
        x = np.abs(np.asarray(x, dtype=float))
        nd = self.nd
        with np.errstate(invalid='ignore', divide='ignore'):
            s = np.sin(np.pi * x)
            sn = np.sin(np.pi * x / nd)
            res = (nd / (np.pi * np.pi)) * s * sn / (x * x)
        pix = np.pi * x
        temp = (1. / 6.) * pix * pix
        small = x <= 1.e-4
        s = np.where(small, pix * (1. - temp), s)
        res = np.where(small, 1. - temp * (1. + 1. / (nd * nd)), res)
        if self.conserve_dc:
            K = self.K
            ssq = s * s
            factor = (1. - 4. * K[1] * ssq - 16. * K[2] * ssq * (1. - ssq)
                      - 4. * K[3] * ssq * (9. - ssq * (24. - 16. * ssq))
                      - 64. * K[4] * ssq * (1. - ssq * (5. - ssq * (8. - 4. * ssq)))
                      - 4. * K[5] * ssq * (25. - ssq * (200. - ssq * (560. - ssq * (640. - 256. *
                                                                                   ssq)))))
            res = res / factor
        return np.where(x >= nd, 0., res)

    def uval(self, u):
        u = np.abs(np.asarray(u, dtype=float))
        return np.where(u > self.umax, 0., self.utab(u))


def interpolant_from_name(name, gsparams=None):
    """Interpolant.from_name (galsim/interpolant.py:45-110)."""
    if not isinstance(name, str):
        return name
    low = name.lower()
    if low == 'quintic':
        return Quintic(gsparams)
    if low.startswith('lanczos'):
        dc = True
        if name[-1].upper() in ('T', 'F'):
            dc = name[-1].upper() == 'T'
            name = name[:-1]
        return Lanczos(int(name[7:]), dc, gsparams)
    if low == 'linear':
        return Linear(gsparams)
    if low == 'cubic':
        return Cubic(gsparams)
    if low == 'nearest':
        return Nearest(gsparams)
    raise ValueError("interpolant %r is not restated" % name)


# ----------------------------------------------------------------------------------------------
# SBInterpolatedImage
# ----------------------------------------------------------------------------------------------
def wrap_k_index(k, No2, N):
    """WrapKIndex (src/SBInterpolatedImage.cpp:217-222)."""
    k = (k + No2) % N
    return k - No2


class SBInterpolatedImage:
    """src/SBInterpolatedImage.cpp.  ``xim`` is the padded image on (-N/2..N/2-1)^2 (array
    [y + N/2, x + N/2]); ``init_bounds`` / ``nonzero_bounds`` are (xmin, xmax, ymin, ymax)."""

    def __init__(self, xim, init_bounds, nonzero_bounds, xinterp, kinterp, stepk, maxk,
                 gsparams):
        self.xim = np.asarray(xim, dtype=float)
        self.N = self.xim.shape[1]
        self.init_bounds = init_bounds
        self.nonzero_bounds = nonzero_bounds
        self.xinterp = xinterp
        self.kinterp = kinterp
        self.gsparams = gsparams
        if stepk <= 0.:
            xmin, xmax, ymin, ymax = init_bounds
            R = max((xmax - xmin) / 2., (ymax - ymin) / 2.)
            R2 = xinterp.xrange
            R = math.sqrt(R * R + R2 * R2)
            stepk = math.pi / R
        self.stepk = stepk
        self.maxk1 = xinterp.urange / USCALE
        self.maxk = self.maxk1 if maxk <= 0. else maxk
        self._kimage = None

    def _sub(self, b):
        xmin, xmax, ymin, ymax = b
        h = self.N // 2
        return self.xim[ymin + h: ymax + h + 1, xmin + h: xmax + h + 1]

    @property
    def flux(self):
        """getFlux (:1064-1078): the sum over the nonzero image."""
        return float(np.sum(self._sub(self.nonzero_bounds), dtype=float))

    @property
    def kimage(self):
        """checkK (:450-464): rfft(shift_in, shift_out) onto (0..N/2, -N/2..N/2-1); array
        [q + N/2, p]."""
        if self._kimage is None:
            self._kimage = rfft2(self.xim, True, True)
        return self._kimage

    def kelem(self, p, q):
        """The stored half-plane read at (p, q), p in [-N/2, N/2], q in [-N/2, N/2-1]: for
        p < 0 the conjugate at (-p, -q), with q = -N/2 kept on its own row (its negative is not
        stored) -- KValueInnerLoop (:324-339)."""
        K = self.kimage
        No2 = self.N // 2
        p = np.asarray(p)
        q = np.asarray(q)
        neg = p < 0
        qq = np.where(neg, np.where(q == -No2, q, -q), q)
        v = K[qq + No2, np.abs(p)]
        return np.where(neg, np.conj(v), v)

    def _p_walk(self, p1, n):
        """The wrapped indices KValueInnerLoop visits for p = p1, p1+1, ...: WrapKIndex(p1), then
        +1 each step with p = N/2 + 1 turned into -N/2 + 1 (:322-326)."""
        No2 = self.N // 2
        N = self.N
        p = wrap_k_index(np.asarray(p1)[..., None], No2, N) + np.arange(n)
        return np.where(p >= No2 + 1, p - N, p)

    def _q_walk(self, q1, n):
        """The wrapped indices of kValue's q loop: WrapKIndex(q1), then +1 with q = N/2 turned
        into -N/2 (:438-439)."""
        No2 = self.N // 2
        N = self.N
        q = wrap_k_index(np.asarray(q1)[..., None], No2, N) + np.arange(n)
        return np.where(q >= No2, q - N, q)

    def _taps(self, k, integer_check):
        """(p1, ntaps) per k value (already scaled by N/(2 pi)): one tap at an integer k when
        integer_check, else ceil(k - xr) .. floor(k + xr)."""
        xr = self.kinterp.xrange
        k = np.asarray(k, dtype=float)
        p1 = np.ceil(k - xr).astype(np.int64)
        p2 = np.floor(k + xr).astype(np.int64)
        if integer_check:
            isint = np.abs(k - np.floor(k + 0.01)) < SMALL * (np.abs(k) + 1.)
            fl = np.floor(k + 0.01).astype(np.int64)
            p1 = np.where(isint, fl, p1)
            p2 = np.where(isint, fl, p2)
        return p1, p2 - p1 + 1

    def _weights(self, k, p1, n, T):
        """kInterp.xval(p - k) for taps p = p1 .. p1 + n - 1 (0 past a row's own taps)."""
        t = np.arange(T)
        p = p1[..., None] + t
        w = self.kinterp.xval(p - np.asarray(k)[..., None])
        return np.where(t < n[..., None], w, 0.)

    def _interp(self, kx, ky, integer_check):
        """sum_q ywt(q) sum_p xwt(p) K(p, q) at arrays kx, ky (scaled, broadcastable)."""
        kx, ky = np.broadcast_arrays(np.asarray(kx, float), np.asarray(ky, float))
        p1, npx = self._taps(kx, integer_check)
        q1, nqy = self._taps(ky, integer_check)
        T = int(max(npx.max(initial=1), nqy.max(initial=1)))
        wx = self._weights(kx, p1, npx, T)
        wy = self._weights(ky, q1, nqy, T)
        P = self._p_walk(p1, T)
        Q = self._q_walk(q1, T)
        vals = self.kelem(P[..., None, :], Q[..., :, None])     # [..., q tap, p tap]
        return np.einsum('...qp,...p,...q->...', vals, wx, wy)

    def kValue(self, kx, ky):
        """kValue (:385-448)."""
        if abs(kx) > self.maxk1 or abs(ky) > self.maxk1:
            return 0j
        xkt = float(self.xinterp.uval(kx * USCALE) * self.xinterp.uval(ky * USCALE))
        No2 = self.N // 2
        kscale = No2 / math.pi
        s = self._interp(np.array(kx * kscale), np.array(ky * kscale), True)
        return complex(s) * xkt

    def fill_k_sep(self, kx0, dkx, ky0, dky, m, n):
        """fillKImage(im, kx0, dkx, izero, ky0, dky, jzero) (:767-936) onto an n x m image: zero
        outside columns i1..i2-1 and rows j1..j2-1, where |k| <= maxk1; the k values advance by
        repeated addition, as the reference's loops do."""
        out = np.zeros((n, m), dtype=np.complex128)
        absdkx, absdky = abs(dkx), abs(dky)
        i1 = max(int(-self.maxk1 / absdkx - kx0 / dkx), 0)
        i2 = min(int(self.maxk1 / absdkx - kx0 / dkx) + 1, m)
        j1 = max(int(-self.maxk1 / absdky - ky0 / dky), 0)
        j2 = min(int(self.maxk1 / absdky - ky0 / dky) + 1, n)
        if i1 >= m or i2 < 0 or j1 >= n or j2 < 0 or i1 >= i2 or j1 >= j2:
            return out
        kx0 += i1 * dkx
        ky0 += j1 * dky
        ux = _accumulate(kx0, dkx, i2 - i1) * USCALE
        uy = _accumulate(ky0, dky, j2 - j1) * USCALE
        No2 = self.N // 2
        kscale = No2 / math.pi
        kx = _accumulate(kx0 * kscale, dkx * kscale, i2 - i1)
        ky = _accumulate(ky0 * kscale, dky * kscale, j2 - j1)
        s = self._interp(kx[None, :], ky[:, None], True)
        out[j1:j2, i1:i2] = self.xinterp.uval(ux)[None, :] * self.xinterp.uval(uy)[:, None] * s
        return out

    def fill_k_gen(self, kx0, dkx, dkxy, ky0, dky, dkyx, m, n):
        """fillKImage(im, kx0, dkx, dkxy, ky0, dky, dkyx) (:938-1009): per sample, no integer
        shortcut, zero where |kx| or |ky| (scaled) exceeds maxk1 (scaled); positions advance by
        repeated addition along each row."""
        No2 = self.N // 2
        kscale = No2 / math.pi
        ux0, uy0 = kx0 * USCALE, ky0 * USCALE
        dux, duy, duxy, duyx = dkx * USCALE, dky * USCALE, dkxy * USCALE, dkyx * USCALE
        kx0, dkx, dkxy = kx0 * kscale, dkx * kscale, dkxy * kscale
        ky0, dky, dkyx = ky0 * kscale, dky * kscale, dkyx * kscale
        maxk1 = self.maxk1 * kscale
        KX = np.empty((n, m))
        KY = np.empty((n, m))
        UX = np.empty((n, m))
        UY = np.empty((n, m))
        for j in range(n):
            KX[j] = _accumulate(kx0, dkx, m)
            KY[j] = _accumulate(ky0, dkyx, m)
            UX[j] = _accumulate(ux0, dux, m)
            UY[j] = _accumulate(uy0, duyx, m)
            kx0 += dkxy
            ky0 += dky
            ux0 += duxy
            uy0 += duy
        cut = (np.abs(KX) > maxk1) | (np.abs(KY) > maxk1)
        s = self._interp(np.where(cut, 0., KX), np.where(cut, 0., KY), False)
        v = self.xinterp.uval(UX) * self.xinterp.uval(UY) * s
        return np.where(cut, 0j, v)

    def calculate_maxk(self, max_maxk=0.):
        """calculateMaxK (:1146-1217)."""
        K = self.kimage
        No2 = self.N // 2
        dk = math.pi / No2

        def norm(p, q):
            z = K[q + No2, p]
            return z.real * z.real + z.imag * z.imag

        thresh = self.gsparams.maxk_threshold * self.flux
        thresh *= thresh
        maxk_ix = 0.
        n_below = 0
        if max_maxk == 0.:
            max_maxk = self.maxk
        max_ix = int(math.ceil(max_maxk / dk))
        if max_ix > No2:
            max_ix = No2
        for ix in range(max_ix + 1):
            for iy in range(ix + 1):
                nk = norm(iy, -ix)
                if nk <= thresh and iy != ix and ix != No2:
                    nk = norm(iy, ix)
                if nk <= thresh and iy > 0:
                    nk = norm(ix, -iy)
                if nk <= thresh and ix > 0 and iy != No2:
                    nk = norm(ix, iy)
                if nk > thresh:
                    maxk_ix = ix
                    n_below = 0
                    break
            n_below += 1
            if n_below == 5:
                break
        maxk_ix += 1
        self.maxk = maxk_ix * dk
        return self.maxk


def _accumulate(x0, dx, n):
    """x0, x0 + dx, (x0 + dx) + dx, ... (n values), as a loop's running sum."""
    out = np.empty(n)
    x = x0
    for i in range(n):
        out[i] = x
        x += dx
    return out


def calculate_size_containing_flux(im, xmin, ymin, target):
    """CalculateSizeContainingFlux (src/SBInterpolatedImage.cpp:1084-1110) on image ``im``
    (array [y - ymin, x - xmin]) whose bounds contain (0, 0)."""
    ny, nx = im.shape
    xmax, ymax = xmin + nx - 1, ymin + ny - 1

    def at(x, y):
        return im[y - ymin, x - xmin]

    p = 1. if target > 0. else -1.
    dmax = min((xmax - xmin) // 2, (ymax - ymin) // 2)
    flux = at(0, 0)
    d = 1
    while d <= dmax:
        for x in range(-d, d):
            flux += at(x, -d)
            flux += at(d, x)
            flux += at(-x, d)
            flux += at(-d, -x)
        if p * flux >= p * target:
            break
        d += 1
    return d + 0.5


# ----------------------------------------------------------------------------------------------
# galsim.InterpolatedImage and its drawing
# ----------------------------------------------------------------------------------------------
class InterpolatedImage(dr.Obj):
    """galsim.InterpolatedImage for an image of pixel scale ``scale`` (a PixelScale wcs): array
    [y, x] with bounds starting at (xmin, ymin).  No noise padding, pad image or depixelize."""

    def __init__(self, array, scale, x_interpolant=None, k_interpolant=None,
                 normalization='flux', flux=None, pad_factor=4., calculate_stepk=True,
                 calculate_maxk=True, use_true_center=True, offset=None, gsparams=None,
                 xmin=1, ymin=1, _force_stepk=0., _force_maxk=0.):
        from . import profiles as pr
        self.gsparams = gsparams or pr.GSParams()
        self.xinterp = (Quintic(self.gsparams) if x_interpolant is None
                        else interpolant_from_name(x_interpolant, self.gsparams))
        self.kinterp = (Quintic(self.gsparams) if k_interpolant is None
                        else interpolant_from_name(k_interpolant, self.gsparams))
        arr = np.asarray(array, dtype=float)
        ny, nx = arr.shape
        # setCenter(0, 0): the centre ((xmin + xmax + 1) // 2, ...) moves to the origin
        cx = (xmin + xmin + nx - 1 + 1) // 2
        cy = (ymin + ymin + ny - 1 + 1) // 2
        bx0, by0 = xmin - cx, ymin - cy
        self.image_bounds = (bx0, bx0 + nx - 1, by0, by0 + ny - 1)
        # _adjust_offset (galsim/gsobject.py:1234-1253)
        ox, oy = (0., 0.) if offset is None else (float(offset[0]), float(offset[1]))
        if use_true_center:
            if nx % 2 == 0:
                ox -= 0.5
            if ny % 2 == 0:
                oy -= 0.5
        self.offset = (ox, oy)
        self.scale = float(scale)
        # _buildRealImage (:447-519)
        pad_size = max(arr.shape)
        if pad_factor > 1.:
            pad_size = int(math.ceil(pad_factor * pad_size))
        pad_size = nm.good_fft_size(pad_size)
        self.pad_size = pad_size
        h = pad_size // 2
        xim = np.zeros((pad_size, pad_size))
        xim[by0 + h: by0 + h + ny, bx0 + h: bx0 + h + nx] = arr
        self.xim = xim
        self.image_array = arr
        self.image_flux = float(np.sum(arr, dtype=np.float64))
        # _getFlux (:568-575)
        if flux is None:
            flux = self.image_flux
            if normalization.lower() in ('surface brightness', 'sb'):
                flux *= self.scale * self.scale
        self._flux = flux
        # _getStepK (:577-611)
        if _force_stepk > 0.:
            self._stepk = _force_stepk
        else:
            if calculate_stepk:
                thresh = (1. - self.gsparams.folding_threshold) * self.image_flux
                R = calculate_size_containing_flux(arr, bx0, by0, thresh)
            else:
                R = max(arr.shape) / 2. - 0.5
            R = math.hypot(R, self.xinterp.xrange)
            self._stepk = math.pi / (R * self.scale)
        # _getMaxK (:613-628)
        if _force_maxk > 0.:
            self._maxk = _force_maxk
        elif calculate_maxk:
            sb = self._sbii(self._stepk, 0.)
            self._maxk = sb.calculate_maxk(0.) / self.scale
        else:
            self._maxk = self.xinterp.krange / self.scale
        self.sbii = self._sbii(self._stepk, self._maxk)
        # _sbp (:398-427): shift by -offset, then the PixelScale transform with the flux ratio,
        # composed into one Transformation (galsim/transform.py:604-640) whose _sbp is the
        # SBTransform(SBInterpolatedImage, jac, cen, flux ratio) that draws
        flux_ratio = self._flux / self.image_flux if self._flux != self.image_flux else 1.
        j = None if self.scale == 1. else np.array([[self.scale, 0.], [0., self.scale]])
        inv = 1. / self.scale
        base = dr.Obj()
        base.gsparams = self.gsparams
        shifted = dr._Transform(base, None, (-ox, -oy), 1.)
        self.world = dr._Transform(shifted, j, (0., 0.), inv * inv * flux_ratio)

    def _sbii(self, stepk, maxk):
        return SBInterpolatedImage(self.xim, self.image_bounds, self.image_bounds, self.xinterp,
                                   self.kinterp, stepk * self.scale, maxk * self.scale,
                                   self.gsparams)

    @property
    def maxk(self):
        return self._maxk

    @property
    def stepk(self):
        return self._stepk

    @property
    def flux(self):
        return self._flux

    def kValue(self, kx, ky):
        """The world profile's kValue: SBTransform::kValue of the shifted, scaled image
        (src/SBTransform.cpp: fwdT of k, the phase exp(-i k.cen), times the flux scaling)."""
        w = self.world
        J = w.jac
        kxp = J[0, 0] * kx + J[1, 0] * ky
        kyp = J[0, 1] * kx + J[1, 1] * ky
        v = self.sbii.kValue(kxp, kyp)
        ph = complex(math.cos(kx * w._dx + ky * w._dy), -math.sin(kx * w._dx + ky * w._dy))
        return v * ph * w._flux_scaling

    def _drawKImage(self, image, jac):
        """InterpolatedImage._drawKImage (galsim/interpolatedimage.py:733-735): _sbp.drawK, i.e.
        SBProfile::drawK (src/SBProfile.cpp:383-425) on the SBTransform, whose fillKImage
        (src/SBTransform.cpp:885-969) hands the SBInterpolatedImage its k grid -- the separable
        form while every Jacobian on the way is diagonal, else the form with cross terms -- and
        then applies the shift's phases or the flux scaling."""
        sb = self.sbii
        w = self.world
        J = w.jac
        A, B, C, D = J[0, 0], J[0, 1], J[1, 0], J[1, 1]
        dk = image.dk
        n, m = image.array.shape
        x0 = image.xmin * dk
        y0 = image.ymin * dk

        def fwdT(kx, ky):
            return A * kx + C * ky, B * kx + D * ky

        if jac is None or (jac[0, 1] == 0. and jac[1, 0] == 0.):
            if jac is None:
                kx0, dkx, ky0, dky = x0, dk, y0, dk
            else:
                kx0, dkx = x0 * jac[0, 0], dk * jac[0, 0]
                ky0, dky = y0 * jac[1, 1], dk * jac[1, 1]
            if B == 0. and C == 0.:
                vals = sb.fill_k_sep(A * kx0, A * dkx, D * ky0, D * dky, m, n)
            else:
                f0 = fwdT(kx0, ky0)
                f1 = fwdT(dkx, 0.)
                f2 = fwdT(0., dky)
                vals = sb.fill_k_gen(f0[0], f1[0], f2[0], f0[1], f2[1], f1[1], m, n)
        else:
            mA, mB, mC, mD = jac[0, 0], jac[0, 1], jac[1, 0], jac[1, 1]
            kx0, dkx, dkxy = mA * x0 + mC * y0, mA * dk, mC * dk
            ky0, dky, dkyx = mB * x0 + mD * y0, mD * dk, mB * dk
            f0 = fwdT(kx0, ky0)
            f1 = fwdT(dkx, dkyx)
            f2 = fwdT(dkxy, dky)
            vals = sb.fill_k_gen(f0[0], f1[0], f2[0], f0[1], f2[1], f1[1], m, n)
        image.array[:, :] = vals
        fs = w._flux_scaling
        if w._has_offset:
            dr.apply_kimage_phases(image, jac, w._dx, w._dy, fs)
        elif abs(fs - 1.) > self.gsparams.kvalue_accuracy:
            image.array *= fs
