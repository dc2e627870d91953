"""TopHat and DeltaFunction: the oracle (oracle/profiles.py) pinned by the reference's known
sizes and by the 2-d Fourier transform of a uniform disk integrated numerically."""
import math

import numpy as np
import pytest
import scipy.integrate as spi
import scipy.special as sps

import galsim_amd as g
from oracle import draw as od
from oracle import profiles as pr


def test_known_sizes():
    # tests/test_box.py:275-297 in the reference
    for m in (g, od):
        t = m.TopHat(radius=0.23, flux=17.9)
        assert t.maxk == pytest.approx(593.7252723959091, rel=1e-14)
        assert t.stepk == pytest.approx(13.659098493868665, rel=1e-14)
    d = g.DeltaFunction(flux=2.)
    assert d.maxk == 1.e300 and d.stepk == 1.e300


def test_kvalue_is_the_disk_transform():
    # F(k) = flux / (pi r^2) int_0^r 2 pi rho J0(k rho) d rho, integrated numerically
    t = pr.TopHat(0.7, flux=3.)
    for k in (0.0, 0.005, 0.01, 0.3, 2.0, 7.5, 31.0):
        val, _ = spi.quad(lambda rho: rho * sps.j0(k * rho), 0., 0.7, epsabs=1e-14,
                          epsrel=1e-13, limit=400)
        expect = 3. * 2. * val / 0.7 ** 2
        got = float(t.kvalue2(np.array([(k * 0.7) ** 2]))[0])
        assert got == pytest.approx(expect, rel=1e-9, abs=1e-13), k


def test_delta_function_is_the_identity_of_convolution():
    gal = od.Exponential(scale_radius=0.4, flux=2.)
    a = od.Convolve([gal, od.DeltaFunction(flux=1.5)]).drawImage(32, 32, scale=0.2)
    b = gal.drawImage(32, 32, scale=0.2) * 1.5
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
