"""CPU tests of libgsmi: every symbol of include/gsmi.h is exported, and the host-side profile
numerics (no device needed) agree with the oracle and with the reference's own compiled code."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import _lib
from oracle import draw as od
from oracle import profiles as pr
from oracle import ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "gsmi.h")).read()
    names = sorted(set(re.findall(r"\b(gsmi_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 15
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes signature table covers them all
    assert set(names) <= set(_lib._SIGS)


def test_record_layouts():
    assert _lib.OP_DTYPE.itemsize == 32
    assert _lib.LEAF_DTYPE.itemsize == 128
    assert _lib.STAMP_DTYPE.itemsize == 80
    assert _lib.STAMP_DTYPE.fields['status'][1] == 72
    assert _lib.STAMP_DTYPE.fields['axis_offset'][1] == 56
    assert _lib.LEAF_DTYPE.fields['jac'][1] == 24
    assert _lib.STAMP_DTYPE.fields['out_offset'][1] == 40


def test_good_fft_size():
    L = _lib.lib()
    from oracle import numerics as nm
    for n in range(1, 5000):
        assert L.gsmi_good_fft_size(n) == nm.good_fft_size(n)


@needs_ref
@pytest.mark.parametrize("n", [0.5, 1.0, 1.5, 2.5, 3.1, 4.0, 5.5, 6.0])
def test_sersic_tables_match_reference(n):
    from galsim_amd.sersic import sersic_info
    sc, args, vals, y2 = sersic_info(n, g.GSParams.default)
    maxk, stepk, hlr, _ = ref.sersic_info(n)
    assert sc[0] == pytest.approx(maxk, rel=1e-14)
    assert sc[1] == pytest.approx(stepk, rel=1e-11)
    assert sc[2] == pytest.approx(hlr, rel=1e-13)
    ksq = np.exp(np.linspace(np.log(1e-8), np.log(1e6), 4001))
    out = np.zeros_like(ksq)
    _lib.check(_lib.lib().gsmi_sersic_kvalue(n, 0., None, _lib.ptr(ksq), _lib.ptr(out),
                                             len(ksq)))
    np.testing.assert_allclose(out, ref.sersic_kvalue(n, ksq), rtol=1e-10, atol=2e-15)


def test_sersic_tables_match_oracle():
    from galsim_amd.sersic import sersic_info
    for n in (1.5, 4.0):
        sc, args, vals, y2 = sersic_info(n, g.GSParams.default)
        info = pr.SersicInfo.get(n, pr.DEFAULT_GSPARAMS)
        np.testing.assert_array_equal(args, info.table.args)
        np.testing.assert_allclose(vals, info.table.vals, rtol=1e-11, atol=1e-16)
        np.testing.assert_allclose(y2, info.table.y2, rtol=1e-8, atol=1e-12)


def test_kolmogorov_and_moffat_scalars():
    k = g.Kolmogorov(lam_over_r0=1.5, flux=1.8)
    # tests/test_kolmogorov.py:138-139 in the reference (5 decimals)
    assert abs(k.maxk - 8.644067599028375) < 1e-5
    assert abs(k.stepk - 0.379103528681262) < 1e-5
    for beta in (1.5, 2., 2.5, 3., 3.5, 4., 2.7, 12.9):
        m = g.Moffat(beta, scale_radius=1.3, flux=2.)
        o = pr.Moffat(beta, scale_radius=1.3, flux=2.)
        assert m.maxk == pytest.approx(o.maxk, rel=1e-12)
        assert m.stepk == pytest.approx(o.stepk, rel=1e-14)


@needs_ref
def test_kolmogorov_stepk_against_reference():
    maxk, stepk = ref.kolmogorov_info()
    k = g.Kolmogorov(lam_over_r0=1.)
    assert k.stepk / k._k0 == pytest.approx(stepk, rel=2e-6)


def test_sizes_match_oracle():
    gal = g.Exponential(half_light_radius=1.5, flux=1e4).shear(g1=0.1, g2=-0.05)
    psf = g.Moffat(beta=3, fwhm=0.7)
    ogal = od.Exponential(half_light_radius=1.5, flux=1e4).shear(g1=0.1, g2=-0.05)
    opsf = od.Moffat(beta=3, fwhm=0.7)
    for a, b in ((g.Convolve(gal, psf), od.Convolve([ogal, opsf])),
                 (g.Sersic(3.1, half_light_radius=2.3) + g.Exponential(scale_radius=1.),
                  od.Sersic(3.1, half_light_radius=2.3) + od.Exponential(scale_radius=1.))):
        assert a.maxk == b.maxk
        assert a.stepk == pytest.approx(b.stepk, rel=1e-13)


@pytest.mark.parametrize("jac", [None, np.diag([1.3, 0.7]),
                                 np.array([[1.1, 0.3], [-0.2, 0.9]]) * 20.])
@pytest.mark.parametrize("r0", [0.05, 0.3, 2.0])
def test_exponential_cliprows_match_oracle(jac, r0):
    Nk, dk = 128, 2 * np.pi / 128
    out = np.zeros(2 * (Nk + 1), dtype=np.int32)
    j = None if jac is None else np.ascontiguousarray(jac.reshape(4))
    _lib.check(_lib.lib().gsmi_exponential_cliprows(r0, None, _lib.ptr(j), Nk, dk,
                                                    _lib.ptr(out)))
    # the oracle's fill of the same k-image: zero exactly outside [i1, i2)
    prof = pr.Exponential(scale_radius=r0)
    img = od.KImage(0, Nk // 2, -Nk // 2, Nk // 2, dk)
    od.draw_k(prof, img, jac)
    mask = img.array != 0
    rows = out.reshape(-1, 2)
    for jrow in range(Nk + 1):
        i1, i2 = rows[jrow]
        expect = np.zeros(Nk // 2 + 1, bool)
        expect[i1:i2] = True
        np.testing.assert_array_equal(mask[jrow], expect)


def test_errors_are_loud():
    with pytest.raises(g.GalSimAMDRangeError):
        g.Sersic(7.0, half_light_radius=1.)
    with pytest.raises(g.GalSimAMDRangeError):
        g.Moffat(1.05, scale_radius=1.)
    with pytest.raises(g.GalSimAMDRangeError):
        g.Sersic(3., half_light_radius=1., trunc=1.4)         # trunc <= sqrt(2) hlr
    out = np.zeros(10)
    rc = _lib.lib().gsmi_sersic_info(float('nan'), 0., None, _lib.ptr(out), None)
    assert rc == -1 and b"bad argument" in _lib.lib().gsmi_last_error()
