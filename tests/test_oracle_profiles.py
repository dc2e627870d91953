"""Pins for the oracle's leaf profiles (oracle/profiles.py): reference known answers, closed
forms that the mathematics fixes, and the reference's own compiled SersicInfo/KolmogorovInfo."""
import math

import numpy as np
import pytest
import scipy.special as sps

from oracle import draw as od
from oracle import profiles as pr
from oracle import ref

GSP = pr.DEFAULT_GSPARAMS
needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def test_exponential_known_answers():
    # tests/test_exponential.py:100-101 in the reference (7 decimals)
    e = pr.Exponential(scale_radius=1.3, flux=2.)
    assert abs(e.maxk - 10 / 1.3) < 1e-7
    assert abs(e.stepk - 0.37436747851 / 1.3) < 1e-7


def test_kolmogorov_known_answers():
    # tests/test_kolmogorov.py:138-139 in the reference (5 decimals)
    k = pr.Kolmogorov(lam_over_r0=1.5, flux=1.8)
    assert abs(k.maxk - 8.644067599028375) < 1e-5
    assert abs(k.stepk - 0.379103528681262) < 1e-5


@needs_ref
def test_kolmogorov_info_against_reference():
    maxk, stepk = ref.kolmogorov_info()
    omaxk, ostepk = pr.kolmogorov_info(GSP)
    assert omaxk == maxk
    # The x-space radial table is an Ogata sum that cancels to ~1e-11 of its terms at the last
    # radius; J0 implementations differing at 1e-17 leave ~1e-6 relative there.
    assert abs(ostepk / stepk - 1.) < 2e-6


def test_sersic_known_answers():
    # tests/test_sersic.py:538-558 in the reference (untruncated n=3.1, hlr=2.3; 7 decimals)
    s = pr.Sersic(3.1, half_light_radius=2.3, flux=1.8)
    assert abs(s.r0 - 0.009526646777227634) < 1e-7
    assert abs(s.maxk - 50.35677579928042) < 1e-7 * 50
    assert abs(s.stepk - 0.08358929429681608) < 1e-7


@needs_ref
@pytest.mark.parametrize("n", [0.5, 1.0, 1.5, 2.5, 3.1, 4.0, 5.5])
def test_sersic_info_against_reference(n):
    info = pr.SersicInfo(n, GSP)
    maxk, stepk, hlr, xnorm = ref.sersic_info(n)
    assert info.maxk == pytest.approx(maxk, rel=1e-13)
    assert info.stepk == pytest.approx(stepk, rel=1e-11)
    assert info.hlr == pytest.approx(hlr, rel=1e-12)
    ksq = np.exp(np.linspace(np.log(1e-8), np.log(1e6), 6001))
    # kValue(0) == 1; 1e-14 absolute is Hankel-quadrature noise on the tiny high-k values
    np.testing.assert_allclose(info.kvalue(ksq), ref.sersic_kvalue(n, ksq), rtol=2e-11,
                               atol=1e-14)


@needs_ref
def test_sersic_hlr_against_reference():
    for n in (0.5, 1., 2.2, 4., 6.):
        assert pr.sersic_hlr(n) == pytest.approx(ref.sersic_hlr(n, 1.), rel=1e-13)


def test_sersic_half_is_exponential_and_gaussian():
    # n=1 is an exponential: kValue(k r0) == (1+k^2 r0^2)^-1.5; n=0.5 a Gaussian.
    info = pr.SersicInfo(1.0, GSP)
    ksq = np.exp(np.linspace(np.log(1e-4), np.log(1e3), 400))
    # the reference's accuracy contract is kvalue_accuracy = 1e-5 of the flux, except just past
    # ksq_min where the natural-spline end condition of the first knots costs ~2e-4 (the
    # reference's own table does the same: see test_sersic_info_against_reference)
    inner = ksq > 1.3 * info.ksq_min
    np.testing.assert_allclose(info.kvalue(ksq)[inner], (1. + ksq[inner]) ** -1.5, rtol=0,
                               atol=1e-5)
    np.testing.assert_allclose(info.kvalue(ksq), (1. + ksq) ** -1.5, rtol=0, atol=2e-4)
    info = pr.SersicInfo(0.5, GSP)
    # exp(-r^2) has transform exp(-k^2/4)
    ksq = np.exp(np.linspace(np.log(1e-4), np.log(30.), 300))
    inner = ksq > 1.3 * info.ksq_min
    np.testing.assert_allclose(info.kvalue(ksq)[inner], np.exp(-ksq[inner] / 4.), rtol=0,
                               atol=2e-5)
    np.testing.assert_allclose(info.kvalue(ksq), np.exp(-ksq / 4.), rtol=0, atol=2e-4)


@pytest.mark.parametrize("beta", [1.5, 1.5001, 2., 2.5, 3., 3.5, 4., 1.22, 3.6, 12.9])
@pytest.mark.parametrize("thresh", [1e-3, 1e-4, 0.03])
def test_moffat_maxk_threshold(beta, thresh):
    # tests/test_moffat.py:183-214 in the reference: kValue(maxk)/flux == maxk_threshold
    gsp = pr.GSParams(maxk_threshold=thresh)
    m = pr.Moffat(beta, scale_radius=1.7, flux=23., gsparams=gsp)
    kv = float(m.knorm * m.kV((m.maxk * m.rD) ** 2)) / m.flux
    assert abs(kv / thresh - 1.) < 1e-7


@pytest.mark.parametrize("beta", [1.5, 2., 2.5, 3., 3.5, 4., 2.7])
def test_moffat_kv_closed_form(beta):
    # The Hankel transform of (1+r^2)^-beta is 2^(1-beta)/Gamma(beta) k^(beta-1) K_{beta-1}(k)
    # times 2 pi (the 2D FT), normalised by the flux pi/(beta-1).
    m = pr.Moffat(beta, scale_radius=1., flux=1.)
    k = np.linspace(0.01, 30., 500)
    expect = 2. ** (2. - beta) * (beta - 1.) / sps.gamma(beta) * k ** (beta - 1.) * \
        sps.kv(beta - 1., k)
    np.testing.assert_allclose(m.knorm * m.kV(k * k), expect, rtol=1e-12, atol=1e-300)


@needs_ref
def test_moffat_hlr_against_reference():
    for beta in (1.5, 2.5, 3., 4.7):
        m = pr.Moffat(beta, half_light_radius=0.9)
        assert m.rD == pytest.approx(ref.moffat_scale_radius_from_hlr(0.9, 0., beta), rel=1e-14)


def test_exponential_drawk_is_moffat_15():
    # tests/test_draw.py:859-897: Exponential(r0).drawKImage == Moffat(beta=1.5, 1/r0).xValue
    flux, r0 = 4.1, 13.
    gal = od.Exponential(scale_radius=r0, flux=flux)
    for scale in (0.15 / r0, 0.6 / r0):
        karr = gal.drawKImage(45, 45, scale)
        x = (np.arange(45) - 22) * scale
        X, Y = np.meshgrid(x, x)
        moff = flux * (1. + (X * X + Y * Y) * r0 * r0) ** -1.5
        np.testing.assert_allclose(karr.real, moff, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(karr.imag, 0.)


def _kvalue(prof, kx, ky):
    """Direct closed-form k-value for a leaf (independent of the fills)."""
    ksq = kx * kx + ky * ky
    if isinstance(prof, pr.Exponential):
        return prof.flux / (1. + ksq * prof.r0 ** 2) ** 1.5
    if isinstance(prof, pr.Gaussian):
        return prof.flux * np.exp(-0.5 * ksq * prof.sigma ** 2)
    if isinstance(prof, pr.Kolmogorov):
        return prof.flux * np.exp(-(ksq / prof.k0 ** 2) ** (5. / 6.))
    if isinstance(prof, pr.Moffat):
        b = prof.beta
        k = np.sqrt(ksq) * prof.rD
        with np.errstate(invalid='ignore'):
            v = 2. ** (2. - b) * (b - 1.) / sps.gamma(b) * k ** (b - 1.) * sps.kv(b - 1., k)
        return prof.flux * np.where(k == 0, 1., v)
    if isinstance(prof, pr.Box):
        return prof.flux * np.sinc(kx * prof.width / (2 * np.pi)) * \
            np.sinc(ky * prof.height / (2 * np.pi))
    raise TypeError


@pytest.mark.parametrize("mk", [
    lambda: pr.Exponential(scale_radius=1.2, flux=3.),
    lambda: pr.Gaussian(sigma=0.8, flux=2.),
    lambda: pr.Kolmogorov(lam_over_r0=0.7, flux=1.5),
    lambda: pr.Moffat(3., scale_radius=0.9, flux=1.1),
    lambda: pr.Moffat(2.7, scale_radius=0.9, flux=1.1),
    lambda: pr.Box(1.3, 0.7, flux=2.),
])
def test_fills_match_closed_form(mk):
    prof = mk()
    dk = 0.11
    m, n = 17, 23
    # general Jacobian
    J = np.array([[1.1, 0.3], [-0.2, 0.9]])
    x0, y0 = 0.0, -11 * dk
    f2 = prof.fill_2d(J[0, 0] * x0 + J[1, 0] * y0, J[0, 0] * dk, J[1, 0] * dk,
                      J[0, 1] * x0 + J[1, 1] * y0, J[1, 1] * dk, J[0, 1] * dk, m, n)
    kx = x0 + np.arange(m) * dk
    ky = y0 + np.arange(n) * dk
    KX, KY = np.meshgrid(kx, ky)
    kxp = J[0, 0] * KX + J[1, 0] * KY
    kyp = J[0, 1] * KX + J[1, 1] * KY
    expect = _kvalue(prof, kxp, kyp)
    # Exponential and Gaussian clip to 0 where the value drops below kvalue_accuracy*flux
    tol = 1.01e-5 * prof.flux if isinstance(prof, (pr.Exponential, pr.Gaussian)) else 1e-12
    np.testing.assert_allclose(f2, expect, rtol=0, atol=tol)
    # diagonal (1-d fill)
    f1 = prof.fill_1d(x0 * 1.3, dk * 1.3, m, y0 * 0.8, dk * 0.8, n)
    np.testing.assert_allclose(f1, _kvalue(prof, 1.3 * KX, 0.8 * KY), rtol=0, atol=tol)


def test_exponential_clipping_only_beyond_kmax():
    # values are zero exactly where k^2 r0^2 > kvalue_accuracy^(-2/3) - 1 (to the sample)
    prof = pr.Exponential(scale_radius=1.0, flux=1.)
    f = prof.fill_1d(0., 1.7, 40, -20 * 1.7, 1.7, 41)
    kx = np.arange(40) * 1.7
    ky = (np.arange(41) - 20) * 1.7
    KX, KY = np.meshgrid(kx, ky)
    ksq = KX ** 2 + KY ** 2
    inside = ksq < prof.ksq_max * (1 - 1e-9)
    outside = ksq > prof.ksq_max * (1 + 1e-9)
    assert np.all(f[inside] > 0) and np.all(f[outside] == 0)
