"""Truncated Moffat (src/SBMoffat.cpp: setupFT 565-626, kV_trunc 340-345, maxK/stepK 497-548,
MoffatCalculateScaleRadiusFromHLR 127-161) on the host: the product's kV_trunc table against the
oracle's (scipy QUADPACK) and the reference's own hankel_trunc (oracle/_ref), and the sizes
against the known answers of tests/test_moffat.py."""
import math

import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import _lib
from galsim_amd.gsparams import GSParams
from oracle import profiles as op
from oracle import ref


CASES = [(3.0, 0.7, 2.0), (2.5, 1.0, 5.0), (1.05, 0.5, 3.0), (4.3, 0.3, 1.1)]


def _product_table(beta, rD, trunc, gsp=None):
    ga = _lib.ptr(_lib.gsp_array(gsp or GSParams()))
    info = np.zeros(6)
    _lib.check(_lib.lib().gsmi_moffat_trunc_info(beta, rD, trunc, ga, _lib.ptr(info)), "info")
    n = int(info[4])
    a, v, y2 = np.zeros(n), np.zeros(n), np.zeros(n)
    _lib.check(_lib.lib().gsmi_moffat_trunc_table(beta, rD, trunc, ga, _lib.ptr(a), _lib.ptr(v),
                                                   _lib.ptr(y2), n), "table")
    return info, a, v, y2


@pytest.mark.parametrize("beta,rD,trunc", CASES)
def test_table_matches_oracle_and_reference_hankel(beta, rD, trunc):
    info, a, v, y2 = _product_table(beta, rD, trunc)
    o = op.Moffat(beta, scale_radius=rD, trunc=trunc)
    # same knots (the same running sum), same values to the integrators' accuracy
    assert np.array_equal(a, o.ft.args)
    assert np.max(np.abs(v - o.ft.vals)) < 1e-12
    # the splines between the knots (second derivatives from divided differences of values
    # 1e-3 apart in k^2 near k = 0 carry the value differences amplified; the spline does not)
    mid = 0.5 * (a[1:] + a[:-1])
    ours = op.nm.SplineTable(a, v)
    ours.y2 = y2
    assert np.max(np.abs(ours(mid) - o.ft(mid))) < 1e-12
    assert abs(info[0] - o.maxk) < 1e-10 * o.maxk and abs(info[1] - o.stepk) < 1e-12
    assert abs(v[0] - 1.) < 1e-13                        # kV_trunc(0) = 1: unit flux
    if ref.available():
        R = trunc / rD
        pre = 2. * (beta - 1.) / (1. - (1. + R * R) ** (1. - beta))
        k = np.sqrt(a[::7])
        rv = pre * np.array([ref.lib().gsref_hankel_moffat_trunc(beta, kk, R, 1e-6, 1e-12)
                             for kk in k])
        assert np.max(np.abs(v[::7] - rv)) < 1e-11


def test_sizes_match_reference_known_answers():
    # tests/test_moffat.py:130-160 (beta = 2, truncated at twice the fwhm, flux 1.8)
    fwhm = 1.4686232496771867
    for m in (g.Moffat(2.0, fwhm=fwhm, trunc=2 * fwhm, flux=1.8),
              op.Moffat(2.0, fwhm=fwhm, trunc=2 * fwhm, flux=1.8)):
        assert abs(m.maxk - 11.634597424960159) < 1e-9
        assert abs(m.stepk - 0.62831853071795873) < 1e-12
        assert abs(m.half_light_radius - 1.0) < 1e-7
    # from the half-light radius: rD by Brent to 1e-7, so hlr (and stepk = pi / 5 hlr) carry
    # ~4e-10; the reference's test checks 7 decimals
    for m in (g.Moffat(2.0, half_light_radius=1., trunc=2 * fwhm, flux=1.8),
              op.Moffat(2.0, half_light_radius=1., trunc=2 * fwhm, flux=1.8)):
        assert abs(m.maxk - 11.634597426100862) < 1e-9
        assert abs(m.stepk - 0.62831853071795862) < 1e-7
    if ref.available():
        rd = ref.moffat_scale_radius_from_hlr(1., 2 * fwhm, 2.0)
        assert g.Moffat(2.0, half_light_radius=1., trunc=2 * fwhm).scale_radius == rd


@pytest.mark.parametrize("beta,rD,trunc", [(1.22, 7., 30.), (3.6, 9., 50.)])
@pytest.mark.parametrize("thresh", [1e-3, 1e-4, 0.03])
def test_kvalue_at_maxk_is_threshold(beta, rD, trunc, thresh):
    """tests/test_moffat.py:183-214: |kValue(maxk)| / flux = maxk_threshold within 3e-3 for
    truncated profiles (maxk comes from linear interpolation between two table knots)."""
    gsp = GSParams(maxk_threshold=thresh)
    m = g.Moffat(beta, scale_radius=rD, trunc=trunc, flux=23., gsparams=gsp)
    info, a, v, y2 = _product_table(beta, rD, trunc, gsp)
    x = (m.maxk * rD) ** 2
    fk = op.nm.SplineTable(a, v)(x)
    assert abs(abs(float(fk)) - thresh) <= 3e-3 * thresh


def test_truncation_argument_checks():
    with pytest.raises(g.GalSimAMDRangeError):
        g.Moffat(3., half_light_radius=1., trunc=1.4)        # trunc <= sqrt(2) hlr
    with pytest.raises(g.GalSimAMDRangeError):
        g.Moffat(3., scale_radius=3., trunc=-1.)
    with pytest.raises(g.GalSimAMDRangeError):
        g.Moffat(1.05, scale_radius=1.)                      # beta <= 1.1 untruncated
    m = g.Moffat(1.05, scale_radius=1., trunc=4.)            # beta <= 1.1 truncated is fine
    assert m.stepk == pytest.approx(math.pi / 4.)
    # hlr input with truncation: the oracle's Brent gives the same radius
    assert g.Moffat(3.7, half_light_radius=2.3, trunc=8.2).scale_radius == pytest.approx(
        op.Moffat(3.7, half_light_radius=2.3, trunc=8.2).rD, rel=1e-14)
