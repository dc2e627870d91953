"""galsim_amd: GalSim's FFT rendering path (GSObject.drawImage(method='fft')) on AMD Instinct
MI355X.  The public names mirror galsim's: profiles, transformations, Convolve/Add, Image,
GSParams, Shear; drawing runs in hand-written HIP kernels through the C ABI of libgsmi.so
(include/gsmi.h)."""
from . import fft
from .airy import Airy
from .angle import Angle, AngleUnit, arcmin, arcsec, degrees, hours, radians
from ._lib import (GalSimAMDBoundsError, GalSimAMDError, GalSimAMDRangeError,
                   GalSimAMDUndefinedBoundsError, GalSimAMDValueError)
from .box import Box, DeltaFunction, Pixel, TopHat
from .convolve import (AutoConvolution, AutoConvolve, AutoCorrelate, AutoCorrelation, Convolution,
                       Convolve, Deconvolution, Deconvolve)
from .exponential import Exponential
from .gaussian import Gaussian
from .gsobject import GSObject
from .gsparams import GSParams
from .image import Image, ImageCD, ImageCF, ImageD, ImageF
from .inclined import InclinedExponential, InclinedSersic
from .interpolant import Cubic, Interpolant, Lanczos, Linear, Nearest, Quintic
from .interpolatedimage import InterpolatedImage
from .kolmogorov import Kolmogorov
from .moffat import Moffat
from .position import BoundsI, PositionD, PositionI
from .sersic import DeVaucouleurs, Sersic
from .spergel import Spergel
from .shear import Shear
from .sum import Add, Sum
from .transform import Transform, Transformation, _Transform

GalSimError = GalSimAMDError

__all__ = ['Add', 'Airy', 'Angle', 'AngleUnit', 'AutoConvolution', 'AutoConvolve', 'AutoCorrelate', 'AutoCorrelation', 'BoundsI',
           'Box', 'Convolution', 'Convolve', 'DeVaucouleurs', 'Deconvolution', 'Deconvolve', 'DeltaFunction',
           'Exponential',
           'GSObject', 'GSParams', 'GalSimAMDBoundsError', 'GalSimAMDError', 'GalSimAMDRangeError',
           'GalSimAMDUndefinedBoundsError', 'GalSimAMDValueError', 'GalSimError',
           'Cubic', 'Gaussian', 'Image', 'InclinedExponential', 'InclinedSersic', 'Interpolant',
           'InterpolatedImage', 'Lanczos', 'Linear', 'Nearest', 'Quintic', 'ImageCD', 'ImageCF', 'ImageD', 'ImageF', 'Kolmogorov', 'Moffat',
           'Pixel', 'PositionD', 'PositionI', 'Sersic', 'Shear', 'Spergel', 'Sum', 'TopHat', 'Transform',
           'Transformation', 'arcmin', 'arcsec', 'degrees', 'fft', 'hours', 'radians']
