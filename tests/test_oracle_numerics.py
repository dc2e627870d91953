"""Pins for the oracle's numerics (oracle/numerics.py) against brute force, library routines
and the reference's own compiled numerics (oracle/_ref)."""
import math

import numpy as np
import pytest
import scipy.interpolate
import scipy.optimize

from oracle import numerics as nm
from oracle import ref


def test_good_fft_size_known_answer():
    # tests/test_draw.py:1306-1307 in the reference
    assert nm.good_fft_size(1174) == 1536 == 3 * 2 ** 9


def test_good_fft_size_brute_force():
    cands = sorted({2 ** k for k in range(1, 16)} | {3 * 2 ** k for k in range(1, 15)})
    for n in range(1, 9000):
        expect = 2 if n <= 2 else min(c for c in cands if c >= n)
        assert nm.good_fft_size(n) == expect, n


def test_brent_against_brentq():
    funcs = [(lambda x: x ** 3 - 2. * x - 5., 2., 3.),
             (lambda x: math.exp(-x) - 0.3 * x, 0., 5.),
             (lambda x: math.cos(x) - x, 0., 1.)]
    for f, a, b in funcs:
        s = nm.Solve(f, a, b)
        r = s.root()
        assert abs(r - scipy.optimize.brentq(f, a, b, xtol=1e-14)) < 1e-7


def test_brent_bracket_upper_expands():
    f = lambda x: x - 37.3
    s = nm.Solve(f, 0., 1.)
    s.bracket_upper()
    assert s.lb <= 37.3 <= s.ub
    assert abs(s.root() - 37.3) < 1e-7


def test_brent_bracket_lower_with_limit():
    f = lambda x: x * x - 0.04     # root at 0.2, bracket must stay above 0
    s = nm.Solve(f, 1., 2.)
    s.bracket_lower_with_limit(0.)
    assert 0. < s.lb <= 0.2
    assert abs(s.root() - 0.2) < 1e-7


@pytest.mark.parametrize("equal", [True, False])
def test_spline_against_scipy_natural(equal):
    rng = np.random.default_rng(5)
    if equal:
        x = np.cumsum(np.full(40, 0.0316))
    else:
        x = np.cumsum(rng.uniform(0.01, 0.1, 40))
    y = np.sin(3 * x) + rng.normal(0, 0.01, x.size)
    tab = nm.SplineTable(x, y)
    assert tab.equal_spaced == equal
    cs = scipy.interpolate.CubicSpline(x, y, bc_type='natural')
    xx = np.linspace(x[0], x[-1], 2001)
    np.testing.assert_allclose(tab(xx), cs(xx), rtol=0, atol=1e-12)
    np.testing.assert_allclose(tab.y2, cs(x, 2), rtol=0, atol=1e-10)
    # outside the range the reference returns 0 (src/Table.cpp:780-785)
    assert tab(x[0] - 1e-3) == 0. and tab(x[-1] + 1e-3) == 0.


def test_expd_close_to_exp():
    # no Cody-Waite split in the argument reduction: the error grows like |x| * eps
    x = np.random.default_rng(1).uniform(-700, 700, 100000)
    rel = np.abs(nm.expd(x) / np.exp(x) - 1.)
    assert np.all(rel < 1.5e-16 * np.maximum(np.abs(x), 30.))
    assert nm.expd(-800.) == 0. and nm.expd(800.) == np.inf


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
def test_expd_bit_exact_against_reference():
    L = ref.lib()
    x = np.random.default_rng(2).uniform(-708, 709, 5000)
    got = nm.expd(x)
    exp = np.array([L.gsref_expd(float(v)) for v in x])
    assert np.array_equal(got, exp)


def test_hankel_closed_form():
    # int_0^inf exp(-r) J0(k r) r dr = (1+k^2)^(-3/2)
    f = lambda r: np.exp(-r)
    for k in (0.05, 0.3, 1., 4., 20., 80.):
        got = nm.hankel_inf(f, k, 1e-10, 1e-14)
        assert abs(got / (1. + k * k) ** -1.5 - 1.) < 1e-8, k


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
def test_hankel_against_reference():
    L = ref.lib()
    from oracle.profiles import fast_pow
    for n in (0.7, 1.5, 4.0):
        invn = 1. / n
        f = lambda r: nm.expd(-fast_pow(r, invn))
        for k in (0.01, 0.3, 2., 30.):
            a = nm.hankel_inf(f, k, 1e-6, 1e-8)
            b = L.gsref_hankel_sersic(invn, k, 1e-6, 1e-8)
            assert abs(a - b) <= 1e-9 * abs(b), (n, k, a, b)


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
def test_bessel_zeros_against_reference():
    import scipy.special as sps
    L = ref.lib()
    z = sps.jn_zeros(0, 3000)
    for s in (1, 2, 7, 50, 500, 2999):
        assert abs(L.gsref_bessel_root(0., s) - z[s - 1]) < 1e-12 * z[s - 1]
