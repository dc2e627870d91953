"""Image: a numpy-backed image over integer bounds (galsim/image.py), the subset the FFT
drawing path and its tests use."""
import numpy as np

from .position import BoundsI, PositionI

_VALID = (np.float32, np.float64, np.complex64, np.complex128)


class Image:
    """``Image(ncol, nrow, dtype=..., scale=...)`` or ``Image(bounds=..., ...)`` or
    ``Image(array, xmin=1, ymin=1)``.  ``array[y, x]`` with x fastest, as galsim.Image."""

    def __init__(self, *args, bounds=None, dtype=np.float32, scale=None, xmin=1, ymin=1,
                 array=None, init_value=0):
        if args and isinstance(args[0], np.ndarray):
            array = args[0]
            args = ()
        if array is not None:
            array = np.asarray(array)
            if array.dtype.type not in _VALID:
                array = array.astype(np.float64)
            self._array = array
            ny, nx = array.shape
            self.bounds = BoundsI(xmin, xmin + nx - 1, ymin, ymin + ny - 1)
        else:
            dtype = np.dtype(dtype).type
            if dtype not in _VALID:
                raise ValueError("unsupported dtype %r" % dtype)
            if bounds is None:
                if len(args) == 2:
                    nx, ny = int(args[0]), int(args[1])
                    bounds = BoundsI(xmin, xmin + nx - 1, ymin, ymin + ny - 1)
                elif len(args) == 1 and isinstance(args[0], BoundsI):
                    bounds = args[0]
                else:
                    bounds = BoundsI()
            self.bounds = bounds
            self._array = np.full(bounds.numpyShape(), init_value, dtype=dtype)
        self.scale = scale
        self.added_flux = None

    @property
    def array(self):
        return self._array

    @property
    def dtype(self):
        return self._array.dtype.type

    @property
    def iscomplex(self):
        return self._array.dtype.kind == 'c'

    @property
    def center(self):
        return self.bounds.center

    @property
    def true_center(self):
        return self.bounds.true_center

    @property
    def xmin(self):
        return self.bounds.xmin

    @property
    def ymin(self):
        return self.bounds.ymin

    def _shift(self, delta):
        self.bounds = self.bounds.shift(delta)

    def shift(self, dx, dy=None):
        if dy is None:
            dx, dy = dx.x, dx.y
        self._shift(PositionI(dx, dy))

    def setCenter(self, x, y=None):
        if y is None:
            x, y = x.x, x.y
        c = self.center
        self._shift(PositionI(int(x) - c.x, int(y) - c.y))

    def setOrigin(self, x, y=None):
        if y is None:
            x, y = x.x, x.y
        self._shift(PositionI(int(x) - self.bounds.xmin, int(y) - self.bounds.ymin))

    def copy(self):
        im = Image(self._array.copy(), xmin=self.bounds.xmin, ymin=self.bounds.ymin)
        im.scale = self.scale
        return im

    def subImage(self, bounds):
        b = self.bounds
        sub = self._array[bounds.ymin - b.ymin: bounds.ymax - b.ymin + 1,
                          bounds.xmin - b.xmin: bounds.xmax - b.xmin + 1]
        im = Image(sub, xmin=bounds.xmin, ymin=bounds.ymin)
        im.scale = self.scale
        return im

    def __call__(self, x, y):
        return self._array[y - self.bounds.ymin, x - self.bounds.xmin]

    def calculate_fft(self):
        """FFT of a real-space image into a k-space image on (0..N/2, -N/2..N/2-1), zero-padded
        to (-N/2..N/2-1)^2 first (galsim/image.py:915-956), on the GPU."""
        from . import render
        return render.forward_fft(self)

    def calculate_inverse_fft(self):
        """Inverse FFT of a k-space image on bounds (0..N/2, -N/2..N/2-1), padded and wrapped
        first when its bounds differ (galsim/image.py:958-1016), on the GPU."""
        from . import render
        return render.inverse_fft(self)

    def __repr__(self):
        return 'galsim_amd.Image(bounds=%r, dtype=%s, scale=%r)' % (
            self.bounds, self._array.dtype, self.scale)

    @staticmethod
    def good_fft_size(input_size):
        from ._lib import lib
        return lib().gsmi_good_fft_size(int(input_size))


def ImageF(*args, **kw):
    kw['dtype'] = np.float32
    return Image(*args, **kw)


def ImageD(*args, **kw):
    kw['dtype'] = np.float64
    return Image(*args, **kw)


def ImageCF(*args, **kw):
    kw['dtype'] = np.complex64
    return Image(*args, **kw)


def ImageCD(*args, **kw):
    kw['dtype'] = np.complex128
    return Image(*args, **kw)
