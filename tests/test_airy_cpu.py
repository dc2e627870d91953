"""Airy: the oracle's k-space (oracle/profiles.py airy_kvalue) pinned by geometry -- the lens
area of two unit circles and a direct integral of the overlap of two annuli -- and the
reference's known sizes; the library's scalars against the same."""
import math

import numpy as np
import pytest

import galsim_amd as g
from oracle import draw as od
from oracle import profiles as pr


def _intervals(x, r_in, r_out):
    """y-intervals of {r_in^2 <= x^2 + y^2 <= r_out^2} at abscissa x."""
    if abs(x) >= r_out:
        return []
    yo = math.sqrt(r_out * r_out - x * x)
    if abs(x) >= r_in:
        return [(-yo, yo)]
    yi = math.sqrt(r_in * r_in - x * x)
    return [(-yo, -yi), (yi, yo)]


def _overlap_area(obs, d, n=200000):
    """Area of the intersection of two annuli (radii obs..1) with centres d apart, integrating
    the length of the vertical cross-sections (midpoint rule over x)."""
    h = 2. / n
    total = 0.
    for k in range(n):
        x = -1. + (k + 0.5) * h
        a = _intervals(x, obs, 1.)
        b = _intervals(x - d, obs, 1.)
        for lo1, hi1 in a:
            for lo2, hi2 in b:
                total += max(0., min(hi1, hi2) - max(lo1, lo2))
    return total * h


def test_unobscured_is_the_lens_area():
    d = np.linspace(0., 2.2, 45)
    lens = np.where(d < 2., 2. * np.arccos(np.minimum(d / 2., 1.)) -
                    (d / 2.) * np.sqrt(np.maximum(4. - d * d, 0.)), 0.)
    np.testing.assert_allclose(pr.airy_kvalue(d * d, 0.), lens, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("obs", [0.1, 0.32, 0.6])
def test_obscured_matches_direct_overlap_integral(obs):
    for d in (0.05, 0.3, 0.45, 0.7, 1.0, 1.3, 1.65, 1.95):
        got = float(pr.airy_kvalue(np.array([d * d]), obs)[0])
        assert got == pytest.approx(_overlap_area(obs, d, 20000), abs=2e-5), (obs, d)
    assert float(pr.airy_kvalue(np.array([0.]), obs)[0]) == pytest.approx(
        math.pi * (1. - obs * obs), rel=1e-15)
    assert float(pr.airy_kvalue(np.array([4.0]), obs)[0]) == 0.


def test_planted_mistake_is_caught():
    # dropping the inner-circle term of the annulus decomposition must fail the area check
    obs, d = 0.32, 0.3
    t = np.array([d * d])
    wrong = pr._airy_circle_pair(1., 1., t) - 2. * pr._airy_circles(1., obs, 1., obs * obs, t)
    assert abs(float(wrong[0]) - _overlap_area(obs, d, 20000)) > 1e-2


def test_known_sizes():
    # tests/test_airy.py:226-240 in the reference
    for m in (g, od):
        p = m.Airy(lam_over_diam=1.9, flux=17.9, obscuration=0.32)
        assert p.maxk == pytest.approx(3.306939635357677, rel=1e-14)
        assert p.stepk == pytest.approx(0.027742458082373515, rel=1e-14)
    a = g.Airy(lam_over_diam=0.7)
    assert a.half_light_radius == pytest.approx(0.7 * 0.5348321477242647, rel=1e-15)
    assert a.fwhm == pytest.approx(0.7 * 1.028993969962188, rel=1e-15)
    with pytest.raises(NotImplementedError):
        g.Airy(lam_over_diam=0.7, obscuration=0.1).half_light_radius
    b = g.Airy(lam=700., diam=4.)
    assert b.lam_over_diam == pytest.approx(700e-9 / 4. * 206264.80624709636, rel=1e-14)
    with pytest.raises(ValueError):
        g.Airy(lam_over_diam=1., lam=700., diam=4.)
    with pytest.raises(ValueError):
        g.Airy(lam=700.)
