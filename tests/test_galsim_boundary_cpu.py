"""Host side of the _galsim-shaped boundary (galsim_amd._galsim): constructors with the
reference's argument order (pysrc/*.cpp), image views over caller memory, the lowering, and the
argument checks that run before any device work."""
import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import _galsim as G
from galsim_amd import _lib

GSP = G.GSParams(128, 8192, 5e-3, 5, 1e-3, 1e-5, 1e-5, 1, 1e-4, 1e-6, 1e-6, 1e-8, 1e-5)


def test_gsparams_takes_the_reference_thirteen_arguments():
    with pytest.raises(TypeError):
        G.GSParams(128, 8192)
    assert GSP._p.maximum_fft_size == 8192 and GSP._p.kvalue_accuracy == 1e-5


def test_profile_constructors_follow_reference_argument_order():
    # SBMoffat(beta, r0, trunc, flux, gsp); SBSersic(n, r0, flux, trunc, gsp)
    m = G.SBMoffat(3.0, 0.7, 0., 2.5, GSP)
    ref = g.Moffat(3.0, scale_radius=0.7, flux=2.5)
    assert m.getFlux() == 2.5 and m.maxK() == ref.maxk and m.stepK() == ref.stepk
    s = G.SBSersic(2.5, 0.4, 7., 3.0, GSP)
    refs = g.Sersic(2.5, scale_radius=0.4, flux=7., trunc=3.0)
    assert s.maxK() == refs.maxk and s.stepK() == refs.stepk and s.getFlux() == 7.
    k = G.SBKolmogorov(1.5, 1., GSP)
    # tests/test_kolmogorov.py:138-139 of the reference: maxk, stepk at lam_over_r0 = 1.5
    assert abs(k.maxK() - 8.644067599028375) < 1e-5 * 8.644067599028375
    assert abs(k.stepK() - 0.379103528681262) < 1e-5 * 0.379103528681262
    assert G.SBBox(1., 2., 3., GSP).getFlux() == 3.


def test_transform_and_composites_wrap_library_objects():
    e = G.SBExponential(1.0, 10., GSP)
    jac = np.array([[1.1, 0.2], [0.1, 0.9]])
    t = G.SBTransform(e, jac.ctypes.data, 0.3, -0.1, 2.0, GSP)
    assert isinstance(t._obj, g.Transformation)
    np.testing.assert_array_equal(t._obj._jac, jac)
    assert (t._obj._dx, t._obj._dy, t._obj._flux_ratio) == (0.3, -0.1, 2.0)
    c = G.SBConvolve([t, G.SBMoffat(3., 0.7, 0., 1., GSP)], False, GSP)
    a = G.SBAdd([c, e], GSP)
    assert isinstance(a._obj, g.Sum) and isinstance(c._obj, g.Convolution)
    assert abs(a.getFlux() - (10. * 2.0 * abs(np.linalg.det(jac)) + 10.)) < 1e-9


def test_image_view_addresses_caller_memory():
    arr = np.zeros((6, 10), dtype=np.complex128)
    sub = arr[1:5, 2:9]                                    # a strided subimage
    v = G.ImageViewCD(sub.ctypes.data, 1, arr.strides[0] // 16, G.BoundsI(0, 6, -2, 1))
    v._array()[...] = 1 + 2j
    assert np.all(arr[1:5, 2:9] == 1 + 2j) and arr.sum() == 28 * (1 + 2j)
    with pytest.raises(_lib.GalSimAMDError):
        G.ImageViewD(arr.ctypes.data, -1, 10, G.BoundsI(0, 1, 0, 1))


def test_unsupported_wrap_and_irfft_shapes_raise_before_device_work():
    arr = np.zeros((9, 5), dtype=np.complex128)
    v = G.ImageViewCD(arr.ctypes.data, 1, 5, G.BoundsI(0, 4, -4, 4))
    with pytest.raises(NotImplementedError):
        G.wrapImage(v, G.BoundsI(0, 2, -2, 1), False, False)
    out = np.zeros((4, 6))
    with pytest.raises(_lib.GalSimAMDError):
        G.irfft(v, G.ImageViewD(out.ctypes.data, 1, 6, G.BoundsI(-2, 3, -2, 1)), True, True)


def test_lower_to_gsmi_gives_the_batched_records():
    gal = G.SBTransform(G.SBExponential(0.9, 1e4, GSP), 0, 0., 0., 1., GSP)
    conv = G.SBConvolve([gal, G.SBKolmogorov(0.8, 1., GSP), G.SBBox(1., 1., 1., GSP)], False,
                        GSP)
    rec = G.lower_to_gsmi(conv, 128, 128, 2 * np.pi / 128)
    ops = rec['ops']
    assert rec['op_count'] + rec['axis_count'] == len(ops)
    assert rec['axis_count'] == 1                           # the pixel sinc, hoisted
    assert [int(o['code']) for o in ops[:rec['op_count']]] == [
        _lib.OP_LEAF, _lib.OP_LEAF, _lib.OP_MUL]
    types = [int(l['type']) for l in rec['leaves']]
    assert types[:2] == [_lib.LEAF_EXPONENTIAL, _lib.LEAF_KOLMOGOROV]
    # FP64 lowering on the drawFFT k-image: the Exponential carries its exact clip rows
    assert rec['leaves'][0]['cliprows'] >= 0 and len(rec['cliprows']) == 2 * 129
