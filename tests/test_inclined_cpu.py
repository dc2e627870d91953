"""InclinedExponential: the oracle (oracle/profiles.py) pinned against Lance Miller's
independently computed images (the reference's tests/inclined_exponential_images, compared as
tests/test_inclined.py:101-160 does), the face-on limit and its own size roots; the library's
host scalars against the oracle."""
import math
import os

import numpy as np
import pytest

import galsim_amd as g
from oracle import draw as od
from oracle import profiles as pr
from tests.fitsio import read_fits

# (flux, inclination, scale radius, scale height, position angle): tests/test_inclined.py:38-45
PARAMS = [("1.0", "0.0", "3.0", "0.3", "0.0"), ("10.0", "1.3", "3.0", "0.5", "0.0"),
          ("0.1", "0.2", "3.0", "0.5", "0.0"), ("1.0", "0.01", "3.0", "0.5", "0.0"),
          ("2.1", "1.57", "2.5", "0.2", "7.4"), ("1.0", "0.1", "2.0", "1.0", "-0.2"),
          ("1.0", "0.78", "2.0", "0.5", "-0.2")]


@pytest.mark.parametrize("flux,inc,r0,h0,pa", PARAMS)
def test_oracle_matches_lance_miller_images(golden_dir, flux, inc, r0, h0, pa):
    path = os.path.join(golden_dir, "inclined", "galaxy_%s_%s_%s_%s.fits" % (inc, r0, h0, pa))
    ref = read_fits(path)
    ref = ref * (float(flux) / ref.sum())
    ny, nx = ref.shape
    prof = od.InclinedExponential(float(inc), float(r0), float(h0), flux=float(flux))
    img = prof.rotate(float(pa)).drawImage(nx, ny, scale=1.0, offset=(0.5, 0.5))
    core = (slice(nx // 2 - 2, nx // 2 + 3), slice(ny // 2 - 2, ny // 2 + 3))
    edge_on = math.cos(float(inc)) < 0.01
    rtol, atol = (5e-2, 5e-4) if edge_on else (1e-2, 1e-4)
    np.testing.assert_allclose(img[core], ref[core], rtol=rtol, atol=atol * float(flux))


def test_face_on_is_the_exponential():
    p = pr.InclinedExponential(0., 0.8, 0.1)
    e = pr.Exponential(scale_radius=0.8)
    k = np.linspace(0., 30., 301)
    kv = pr.inclined_exp_kvalue(k, 0.3 * k, p.cosi, p.hs, p.ksq_min, p.ksq_max)
    ksq = k * k * 1.09
    ev = np.where(ksq > e.ksq_max, 0., (1. + ksq) ** -1.5)
    np.testing.assert_allclose(kv, ev, rtol=0, atol=2e-5)
    assert p.maxk == pytest.approx(e.maxk, rel=1e-15)
    assert p.stepk == pytest.approx(e.stepk, rel=1e-15)


@pytest.mark.parametrize("inc", [0.2, 0.9, 1.3, 1.5, 1.565, math.pi / 2])
def test_size_roots_and_library_scalars(inc):
    p = pr.InclinedExponential(inc, 1.1, 0.2)
    gsp = pr.DEFAULT_GSPARAMS
    inf = math.inf
    along = lambda k: max(float(pr.inclined_exp_kvalue(k, 0., p.cosi, p.hs, p.ksq_min, inf)),
                          float(pr.inclined_exp_kvalue(0., k, p.cosi, p.hs, p.ksq_min, inf)))
    assert along(p._maxk) < gsp.maxk_threshold <= along(p._maxk - 2e-7)
    clipk = math.sqrt(p.ksq_max)
    assert along(clipk) < gsp.kvalue_accuracy <= along(clipk - 2e-7)
    q = g.InclinedExponential(inc * g.radians, scale_radius=1.1, scale_height=0.2)
    assert q.maxk == pytest.approx(p.maxk, rel=1e-14)
    assert q.stepk == pytest.approx(p.stepk, rel=1e-15)
    info = q._info()[2]
    assert info[2] == pytest.approx(p.ksq_min, rel=1e-15)
    assert info[3] == pytest.approx(p.ksq_max, rel=1e-14)


def test_arguments():
    with pytest.raises(TypeError):
        g.InclinedExponential(0.3, half_light_radius=1.)
    with pytest.raises(ValueError):
        g.InclinedExponential(0.3 * g.radians)
    with pytest.raises(ValueError):
        g.InclinedExponential(0.3 * g.radians, half_light_radius=1., scale_radius=1.)
    with pytest.raises(ValueError):
        g.InclinedExponential(0.3 * g.radians, scale_radius=-1.)
    p = g.InclinedExponential(30 * g.degrees, half_light_radius=1.)
    assert p.scale_height == pytest.approx(0.1 * p.scale_radius)
    assert p.disk_half_light_radius == pytest.approx(1.)
    assert (30 * g.degrees).rad == pytest.approx(math.pi / 6)


@pytest.mark.parametrize("flux,inc,r0,h0,pa", PARAMS)
def test_oracle_inclined_sersic_n1_matches_lance_miller_images(golden_dir, flux, inc, r0, h0, pa):
    # tests/test_inclined.py:101-160, mode InclinedSersic (n = 1, no truncation)
    path = os.path.join(golden_dir, "inclined", "galaxy_%s_%s_%s_%s.fits" % (inc, r0, h0, pa))
    ref = read_fits(path)
    ref = ref * (float(flux) / ref.sum())
    ny, nx = ref.shape
    prof = od.InclinedSersic(1., float(inc), float(r0), float(h0), flux=float(flux))
    img = prof.rotate(float(pa)).drawImage(nx, ny, scale=1.0, offset=(0.5, 0.5))
    core = (slice(nx // 2 - 2, nx // 2 + 3), slice(ny // 2 - 2, ny // 2 + 3))
    rtol, atol = (5e-2, 5e-4) if math.cos(float(inc)) < 0.01 else (1e-2, 1e-4)
    np.testing.assert_allclose(img[core], ref[core], rtol=rtol, atol=atol * float(flux))


# (flux, n, inclination, scale radius, scale height, truncation / r0, position angle):
# tests/test_inclined.py:58-63
SERSIC_PARAMS = [("1.0", "0.5", "0.7", "4.0", "0.2", "0.0", "0.5"),
                 ("1.0", "0.8", "1.4", "2.5", "0.1", "4.0", "1.1"),
                 ("1.0", "1.2", "2.0", "1.5", "0.5", "6.5", "5.5"),
                 ("1.0", "1.5", "0.3", "3.0", "0.3", "0.0", "0.0")]


@pytest.mark.parametrize("flux,n,inc,r0,h0,tf,pa", SERSIC_PARAMS)
def test_oracle_inclined_sersic_matches_lance_miller_images(golden_dir, flux, n, inc, r0, h0, tf,
                                                            pa):
    path = os.path.join(golden_dir, "inclined",
                        "galaxy_%s_%s_%s_%s_%s_%s.fits" % (n, inc, r0, h0, tf, pa))
    ref = read_fits(path)
    ref = ref * (float(flux) / ref.sum())
    ny, nx = ref.shape
    prof = od.InclinedSersic(float(n), float(inc), float(r0), float(h0), flux=float(flux),
                             trunc=float(tf) * float(r0))
    img = prof.rotate(float(pa)).drawImage(nx, ny, scale=1.0, offset=(0.5, 0.5))
    core = (slice(nx // 2 - 2, nx // 2 + 3), slice(ny // 2 - 2, ny // 2 + 3))
    np.testing.assert_allclose(img[core], ref[core], rtol=1e-2, atol=1e-4 * float(flux))


@pytest.mark.parametrize("inc", [0.0, 0.4, 1.2, 1.57])
def test_inclined_sersic_n1_is_inclined_exponential_and_library_scalars(inc):
    s = pr.InclinedSersic(1., inc, 0.9, 0.15)
    e = pr.InclinedExponential(inc, 0.9, 0.15)
    k = np.linspace(0., 20., 201)
    for a, b in ((k, 0.3 * k), (0.2 * k, k)):
        sv = s.helper(a, b)
        ev = pr.inclined_exp_kvalue(a, b, e.cosi, e.hs, e.ksq_min, math.inf)
        np.testing.assert_allclose(sv, ev, rtol=0, atol=3e-5)
    q = g.InclinedSersic(1., inc * g.radians, scale_radius=0.9, scale_height=0.15)
    assert q.maxk == pytest.approx(s.maxk, rel=1e-12)
    assert q.stepk == pytest.approx(s.stepk, rel=1e-13)
    cosi, hs, info = q._info()
    assert info[2] == pytest.approx(s.ksq_max, rel=1e-12)
    if inc > 0:
        assert info[1] == pytest.approx(s.ksq_min, rel=1e-14)
