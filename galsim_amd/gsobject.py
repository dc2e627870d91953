"""GSObject: the base class of all profiles, with drawImage(method='fft') and drawKImage
(galsim/gsobject.py).  Drawing lowers the profile tree into a k-space program and runs it on the
GPU through libgsmi (galsim_amd/render.py)."""
import math

import numpy as np

from .gsparams import GSParams
from .image import Image
from .position import BoundsI, PositionD, PositionI


def _real_values(pixels, dtype, add_to_image):
    """drawReal's drawn values and its return value (galsim/gsobject.py:1875-1889): a float32 or
    float64 image that is overwritten receives the values directly; otherwise they are drawn
    into a temporary, float64 for complex128 / int32 / uint32 images and float32 for every other
    type (complex64, and float64 when adding), and the temporary's sum is returned."""
    if dtype in (np.float32, np.float64) and not add_to_image:
        dt = dtype
    elif dtype in (np.complex128, np.int32, np.uint32):
        dt = np.float64
    else:
        dt = np.float32
    vals = pixels.astype(dt)
    return vals, float(vals.sum(dtype=float))


class GSObject:
    """Base class (galsim/gsobject.py:37-2609).  Subclasses define ``_maxk``, ``_stepk``,
    ``_flux`` and ``_lower(ctx, jac)`` (their contribution to the k-space program)."""
    _is_analytic_x = False
    _has_hard_edges = False

    # -- basic properties ----------------------------------------------------------------------
    @property
    def flux(self):
        return self._flux

    @property
    def gsparams(self):
        return self._gsparams

    @property
    def maxk(self):
        return self._maxk

    @property
    def stepk(self):
        return self._stepk

    @property
    def nyquist_scale(self):
        return math.pi / self.maxk

    @property
    def is_analytic_x(self):
        return self._is_analytic_x

    @property
    def has_hard_edges(self):
        return self._has_hard_edges

    def withGSParams(self, gsparams=None, **kwargs):
        if gsparams == self.gsparams and not kwargs:
            return self
        import copy
        ret = copy.copy(self)
        ret._gsparams = GSParams.check(gsparams, self.gsparams, **kwargs)
        ret._reset()
        return ret

    def _reset(self):
        pass

    # -- arithmetic (galsim/gsobject.py:455-500) -----------------------------------------------
    def __add__(self, other):
        from .sum import Add
        return Add([self, other])

    def __sub__(self, other):
        from .sum import Add
        return Add([self, (-1. * other)])

    def __mul__(self, other):
        return self.withScaledFlux(other)

    def __rmul__(self, other):
        return self.__mul__(other)

    def __truediv__(self, other):
        return self * (1. / other)

    def __neg__(self):
        return -1. * self

    # -- transformations (galsim/gsobject.py:775-1094) -----------------------------------------
    def withFlux(self, flux):
        return self.withScaledFlux(flux / self.flux)

    def withScaledFlux(self, flux_ratio):
        from .transform import Transform
        if flux_ratio == 1:
            return self
        return Transform(self, flux_ratio=flux_ratio)

    def expand(self, scale):
        from .transform import Transform
        return Transform(self, jac=[scale, 0., 0., scale])

    def dilate(self, scale):
        from .transform import Transform
        return Transform(self, jac=[scale, 0., 0., scale], flux_ratio=scale ** -2)

    def magnify(self, mu):
        return self.expand(math.sqrt(mu))

    def shear(self, *args, **kwargs):
        from .shear import Shear
        from .transform import Transform
        if len(args) == 1:
            if kwargs:
                raise TypeError("Error, gave both unnamed and named arguments to shear!")
            if not isinstance(args[0], Shear):
                raise TypeError("Error, unnamed argument to shear is not a Shear!")
            shear = args[0]
        elif len(args) > 1:
            raise TypeError("Error, too many unnamed arguments to shear!")
        elif not kwargs:
            raise TypeError("Error, shear argument is required")
        else:
            shear = Shear(**kwargs)
        return Transform(self, shear.getMatrix())

    def _shear(self, shear):
        from .transform import _Transform
        return _Transform(self, shear.getMatrix())

    def lens(self, g1, g2, mu):
        from .shear import Shear
        from .transform import Transform
        return Transform(self, Shear(g1=g1, g2=g2).getMatrix() * math.sqrt(mu))

    def rotate(self, theta):
        """theta in radians (or any object with a ``.rad`` attribute)."""
        from .transform import Transform
        rad = getattr(theta, 'rad', theta)
        s, c = math.sin(rad), math.cos(rad)
        return Transform(self, jac=[c, -s, s, c])

    def transform(self, dudx, dudy, dvdx, dvdy):
        from .transform import Transform
        return Transform(self, jac=[dudx, dudy, dvdx, dvdy])

    def shift(self, *args, **kwargs):
        from .transform import Transform
        if len(args) == 2:
            offset = PositionD(args[0], args[1])
        elif len(args) == 1:
            offset = PositionD(args[0])
        elif 'dx' in kwargs or 'dy' in kwargs:
            offset = PositionD(kwargs.get('dx', 0.), kwargs.get('dy', 0.))
        elif 'offset' in kwargs:
            offset = PositionD(kwargs['offset'])
        else:
            raise TypeError("shift requires dx, dy or an offset")
        return Transform(self, offset=(offset.x, offset.y))

    def _shift(self, dx, dy):
        from .transform import _Transform
        return _Transform(self, offset=(dx, dy))

    # -- sizes ---------------------------------------------------------------------------------
    def getGoodImageSize(self, pixel_scale):
        """galsim/gsobject.py:1900-1925."""
        Nd = 2. * math.pi / (pixel_scale * self.stepk)
        N = int(math.ceil(Nd * (1. - 1.e-12)))
        return 2 * ((N + 1) // 2)

    # -- drawing -------------------------------------------------------------------------------
    def _setup_image(self, image, nx, ny, bounds, add_to_image, dtype, center, odd=False):
        """galsim/gsobject.py:1115-1172."""
        if image is not None:
            if bounds is not None or nx is not None or ny is not None:
                raise ValueError("Cannot provide bounds or nx,ny if image is provided")
            if not image.bounds.isDefined():
                if add_to_image:
                    raise ValueError("Cannot add_to_image if image bounds are not defined")
                N = self.getGoodImageSize(1.0)
                if odd:
                    N += 1
                image.__init__(bounds=BoundsI(1, N, 1, N), dtype=image.dtype, scale=image.scale)
            return image
        if add_to_image:
            raise ValueError("Cannot add_to_image if image is None")
        dtype = np.float32 if dtype is None else dtype
        if bounds is not None:
            if nx is not None or ny is not None:
                raise ValueError("Cannot set both bounds and (nx, ny)")
            return Image(bounds=bounds, dtype=dtype)
        if nx is not None or ny is not None:
            if nx is None or ny is None:
                raise ValueError("Must set either both or neither of nx, ny")
            image = Image(nx, ny, dtype=dtype)
            if center is not None:
                tc = image.true_center
                image._shift(PositionI(int(np.floor(center.x + 0.5 - tc.x)),
                                       int(np.floor(center.y + 0.5 - tc.y))))
            return image
        N = self.getGoodImageSize(1.0)
        if odd:
            N += 1
        image = Image(N, N, dtype=dtype)
        if center is not None:
            image.setCenter(int(np.ceil(center.x)), int(np.ceil(center.y)))
        return image

    @staticmethod
    def _get_new_bounds(image, nx, ny, bounds, center):
        if image is not None and image.bounds.isDefined():
            return image.bounds
        if nx is not None and ny is not None:
            b = BoundsI(1, nx, 1, ny)
            if center is not None:
                tc = b.true_center
                b = b.shift(PositionI(int(np.floor(center.x + 0.5 - tc.x)),
                                      int(np.floor(center.y + 0.5 - tc.y))))
            return b
        if bounds is not None and bounds.isDefined():
            return bounds
        return BoundsI()

    @staticmethod
    def _adjust_offset(new_bounds, offset, center, use_true_center):
        """galsim/gsobject.py:1234-1253."""
        if center is not None:
            if new_bounds.isDefined():
                c = new_bounds.center
                offset = PositionD(offset.x + center.x - c.x, offset.y + center.y - c.y)
            else:
                offset = PositionD(offset.x + center.x - np.ceil(center.x),
                                   offset.y + center.y - np.ceil(center.y))
        elif use_true_center:
            dx, dy = offset.x, offset.y
            shape = new_bounds.numpyShape()
            if shape[1] % 2 == 0:
                dx -= 0.5
            if shape[0] % 2 == 0:
                dy -= 0.5
            offset = PositionD(dx, dy)
        return offset

    def drawImage(self, image=None, nx=None, ny=None, bounds=None, scale=None, dtype=None,
                  method='auto', area=1., exptime=1., gain=1., add_to_image=False, center=None,
                  use_true_center=True, offset=None, precision='f64', setup_only=False):
        """drawImage (galsim/gsobject.py:1283-1844) for the FFT methods.

        method 'auto' and 'fft' convolve by the pixel; 'no_pixel' and 'sb' draw the profile as
        is ('sb' in surface-brightness units): by direct evaluation at the pixel centres when the
        profile is analytic in real space (drawReal, as the reference does), else by FFT.
        Real-space convolution ('real_space') and photon shooting are not part of this library and
        raise NotImplementedError.  ``precision`` selects the FP64 kernels (default, the
        reference's arithmetic) or the FP32 ones of the FFT path; drawReal is FP64.
        """
        from .box import Pixel
        from .convolve import Convolve
        from .transform import _Transform
        from . import render
        if image is not None and not isinstance(image, Image):
            raise TypeError("image is not an Image instance", image)
        if not gain > 0. or not area > 0. or not exptime > 0.:
            raise ValueError("gain, area and exptime must be > 0")
        if method not in ('auto', 'fft', 'real_space', 'phot', 'no_pixel', 'sb'):
            raise ValueError("Invalid method name %r" % method)
        if method in ('real_space', 'phot'):
            raise NotImplementedError("method=%r is not part of galsim_amd (FFT path only)"
                                      % method)
        if scale is None and image is not None and image.scale is not None:
            scale = image.scale
        if scale is None or not scale > 0:
            scale = self.nyquist_scale
        scale = float(scale)
        offset = PositionD(0., 0.) if offset is None else PositionD(offset)
        if center is not None:
            center = PositionD(center)
        new_bounds = self._get_new_bounds(image, nx, ny, bounds, center)
        flux_scale = area * exptime
        if method == 'sb':
            flux_scale /= scale * scale
        if gain != 1:
            flux_scale /= gain
        offset = self._adjust_offset(new_bounds, offset, center, use_true_center)
        # PixelScale._profileToImage (galsim/wcs.py:1550-1558)
        jac = None if scale == 1. else np.array([[1. / scale, 0.], [0., 1. / scale]])
        prof = _Transform(self, jac, (offset.x, offset.y), scale ** 2 * flux_scale)
        if method in ('auto', 'fft'):
            # 'auto' leaves the choice to Convolution (real space for two hard-edged profiles,
            # galsim/gsobject.py:1754-1764), which raises here: real-space convolution is not
            # part of this library
            prof = Convolve(prof, Pixel(scale=1.0, gsparams=self.gsparams),
                            real_space=None if method == 'auto' else False,
                            gsparams=self.gsparams)
        use_real = method in ('no_pixel', 'sb') and prof.is_analytic_x
        image = prof._setup_image(image, nx, ny, bounds, add_to_image, dtype, center)
        image.scale = scale
        if setup_only:
            image.added_flux = 0.
            return image
        c = image.center
        draw_bounds = image.bounds.shift(PositionI(-c.x, -c.y))
        if use_real:
            from . import real
            pixels, added = _real_values(real.draw_real(prof, draw_bounds, 1.0),
                                         image.array.dtype, add_to_image)
        else:
            pixels, added = render.draw_fft(prof, draw_bounds, precision=precision)
        if add_to_image:
            image.array[...] += pixels.astype(image.array.dtype)
        else:
            image.array[...] = pixels.astype(image.array.dtype)
        image.added_flux = added / flux_scale
        return image

    def drawReal(self, image, add_to_image=False):
        """drawReal (galsim/gsobject.py:1846-1889): the profile sampled at the centre of each
        pixel times the pixel area, its origin at the image's (0, 0), evaluated on the GPU.
        Returns the flux drawn (the sum of the drawn pixel values)."""
        from . import real
        if image.scale is None or not image.scale > 0:
            raise ValueError("drawReal requires an image with a PixelScale wcs", image)
        pixels, added = _real_values(real.draw_real(self, image.bounds, float(image.scale)),
                                     image.array.dtype, add_to_image)
        if add_to_image:
            image.array[...] += pixels.astype(image.array.dtype)
        else:
            image.array[...] = pixels.astype(image.array.dtype)
        return added

    def drawFFT(self, image, add_to_image=False, precision='f64'):
        """drawFFT (galsim/gsobject.py:2014-2045): image bounds centred on the profile origin."""
        from . import render
        pixels, added = render.draw_fft(self, image.bounds, precision=precision)
        if add_to_image:
            image.array[...] += pixels.astype(image.array.dtype)
        else:
            image.array[...] = pixels.astype(image.array.dtype)
        return added

    def drawKImage(self, image=None, nx=None, ny=None, bounds=None, scale=None,
                   recenter=True):
        """drawKImage (galsim/gsobject.py:2464-2583) into a complex128 image, on the GPU."""
        from . import render
        from .transform import _Transform
        if image is not None and scale is None:
            scale = image.scale
        dk = self.stepk if (scale is None or not scale > 0) else float(scale)
        if image is not None and image.bounds.isDefined():
            dx = np.pi / (max(image.array.shape) // 2 * dk)
        elif scale is None or not scale > 0:
            dx = self.nyquist_scale
        else:
            dx = self.nyquist_scale * dk / self.stepk
        if image is None or not image.bounds.isDefined():
            j = None if dx == 1. else np.array([[1. / dx, 0.], [0., 1. / dx]])
            real_prof = _Transform(self, j, (0., 0.), dx * dx)
            image = real_prof._setup_image(image, nx, ny, bounds, False, np.complex128, None,
                                           odd=True)
        if recenter and image.center != PositionI(0, 0):
            c = image.center
            image._shift(PositionI(-c.x, -c.y))
        image.scale = dk
        karr = render.draw_kimage(self, image.bounds, dk)
        image.array[...] = karr.astype(image.array.dtype)
        return image

    def kValue(self, *args):
        """k-space value at a point, evaluated by the device program."""
        from . import render
        kpos = PositionD(*args) if len(args) == 2 else PositionD(args[0])
        return render.kvalue(self, kpos.x, kpos.y)
