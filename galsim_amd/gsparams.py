"""GSParams: the accuracy knobs of the FFT path (galsim/gsparams.py:126-269)."""
import numpy as np


class GSParams:
    """Same names, defaults and combination rule as galsim.GSParams (galsim/gsparams.py:126-241).
    """
    _names = ('minimum_fft_size', 'maximum_fft_size', 'folding_threshold', 'stepk_minimum_hlr',
              'maxk_threshold', 'kvalue_accuracy', 'xvalue_accuracy', 'table_spacing',
              'realspace_relerr', 'realspace_abserr', 'integration_relerr', 'integration_abserr',
              'shoot_accuracy')

    def __init__(self, minimum_fft_size=128, maximum_fft_size=8192, folding_threshold=5.e-3,
                 stepk_minimum_hlr=5, maxk_threshold=1.e-3, kvalue_accuracy=1.e-5,
                 xvalue_accuracy=1.e-5, table_spacing=1, realspace_relerr=1.e-4,
                 realspace_abserr=1.e-6, integration_relerr=1.e-6, integration_abserr=1.e-8,
                 shoot_accuracy=1.e-5):
        self._minimum_fft_size = int(minimum_fft_size)
        self._maximum_fft_size = int(maximum_fft_size)
        self._folding_threshold = float(folding_threshold)
        self._stepk_minimum_hlr = float(stepk_minimum_hlr)
        self._maxk_threshold = float(maxk_threshold)
        self._kvalue_accuracy = float(kvalue_accuracy)
        self._xvalue_accuracy = float(xvalue_accuracy)
        self._table_spacing = int(table_spacing)
        self._realspace_relerr = float(realspace_relerr)
        self._realspace_abserr = float(realspace_abserr)
        self._integration_relerr = float(integration_relerr)
        self._integration_abserr = float(integration_abserr)
        self._shoot_accuracy = float(shoot_accuracy)
        # GSParams are immutable: the argument tuple and the float64 array handed to the native
        # library are built once (every profile constructor and drawing call asks for them)
        self._args = tuple(self.__dict__['_' + n] for n in GSParams._names)
        self._f64 = np.array(self._args, dtype=np.float64)
        self._f64.flags.writeable = False

    def __getattr__(self, name):
        if name in GSParams._names:
            return self.__dict__['_' + name]
        raise AttributeError(name)

    def _getinitargs(self):
        return self._args

    def _getinitargs_f(self):
        return [float(v) for v in self._args]

    def _f64_array(self):
        """The 13 parameters as a read-only contiguous float64 array, in the order above."""
        return self._f64

    def withParams(self, **kwargs):
        if not kwargs:
            return self
        d = dict(zip(GSParams._names, self._getinitargs()))
        for k, v in kwargs.items():
            if k not in d:
                raise TypeError("parameter %s is invalid" % k)
            d[k] = v
        return GSParams(**d)

    @staticmethod
    def check(gsparams, default=None, **kwargs):
        if gsparams is None:
            gsparams = default if default is not None else GSParams.default
        elif not isinstance(gsparams, GSParams):
            raise TypeError("Invalid GSParams: %s" % gsparams)
        return gsparams.withParams(**kwargs)

    @staticmethod
    def combine(gsp_list):
        if len(gsp_list) == 1 or all(g == gsp_list[0] for g in gsp_list[1:]):
            return gsp_list[0]
        return GSParams(
            max(g.minimum_fft_size for g in gsp_list), max(g.maximum_fft_size for g in gsp_list),
            min(g.folding_threshold for g in gsp_list), max(g.stepk_minimum_hlr for g in gsp_list),
            min(g.maxk_threshold for g in gsp_list), min(g.kvalue_accuracy for g in gsp_list),
            min(g.xvalue_accuracy for g in gsp_list), min(g.table_spacing for g in gsp_list),
            min(g.realspace_relerr for g in gsp_list), min(g.realspace_abserr for g in gsp_list),
            min(g.integration_relerr for g in gsp_list),
            min(g.integration_abserr for g in gsp_list), min(g.shoot_accuracy for g in gsp_list))

    def __eq__(self, other):
        return self is other or (isinstance(other, GSParams) and
                                 self._getinitargs() == other._getinitargs())

    def __hash__(self):
        return hash(('galsim_amd.GSParams',) + self._getinitargs())

    def __repr__(self):
        return 'galsim_amd.GSParams(%s)' % ', '.join(repr(v) for v in self._getinitargs())


GSParams.default = GSParams()
