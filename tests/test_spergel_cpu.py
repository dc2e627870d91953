"""Spergel: the oracle (oracle/profiles.py) pinned by the reference's known answers and by
nu = 0.5 being the Exponential, and libgsmi's host scalars and clip rows against it."""
import ctypes

import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import _lib
from oracle import draw as od
from oracle import profiles as pr


def test_oracle_known_answers():
    # tests/test_spergel.py:106-190 in the reference (7 decimals there)
    for nu, r0, maxk, stepk in ((0.2, 1.6003703593406464, 11.094091231317272, 0.2731819698773733),
                                (-0.2, 2.1783499913527824, 34.42181161305098,
                                 0.23712261251942787)):
        p = pr.Spergel(nu, half_light_radius=2.3, flux=17.9)
        assert p.r0 == pytest.approx(r0, rel=1e-13)
        assert p.maxk == pytest.approx(maxk, rel=1e-13)
        assert p.stepk == pytest.approx(stepk, rel=1e-13)
    # tests/test_spergel.py:30-83: Mathematica enclosed fluxes and radii at r = 1e-5 r0
    for nu, flux, radius in ((-0.85, 3.06256e-2, 2.3973e-17), (-0.5, 9.99995e-6, 1.00001e-5),
                             (0.0, 6.06443e-10, 1.69047e-3), (0.85, 2.94117e-11, 5.83138e-3),
                             (4.0, 6.25011e-12, 1.26492e-2)):
        assert pr.spergel_integrated_flux(nu, 1.e-5) / flux == pytest.approx(1., abs=1e-4)
        assert pr.spergel_flux_radius(nu, 1.e-5) / radius == pytest.approx(1., abs=1e-4)


def test_oracle_nu_half_is_exponential():
    # (1 + k^2)^(-3/2), the same clipping radius and the same half-light radius
    assert pr.spergel_flux_radius(0.5, 0.5) == pytest.approx(pr.Exponential.hlr_factor,
                                                             rel=1e-14)
    sp = od.Spergel(0.5, scale_radius=0.37, flux=2.)
    ex = od.Exponential(scale_radius=0.37, flux=2.)
    for jac in (None, np.diag([1.3, 0.8]), np.array([[1.1, 0.3], [-0.2, 0.9]]) * 4.):
        a = od.KImage(0, 64, -64, 64, 0.37)
        b = od.KImage(0, 64, -64, 64, 0.37)
        sp._drawKImage(a, jac)
        ex._drawKImage(b, jac)
        np.testing.assert_array_equal(a.array == 0, b.array == 0)
        np.testing.assert_allclose(a.array, b.array, rtol=1e-14, atol=0)


@pytest.mark.parametrize("nu", [-0.85, -0.6, 0.2, 0.5, 2.0, 4.0])
def test_host_scalars_match_oracle(nu):
    from galsim_amd.spergel import spergel_info
    maxk, stepk, hlr, ksq_max = spergel_info(nu, g.GSParams.default)
    p = pr.Spergel(nu, scale_radius=1.)
    assert maxk == pytest.approx(p.maxk, rel=1e-14)
    assert stepk == pytest.approx(p.stepk, rel=1e-12)
    assert hlr == pytest.approx(pr.spergel_flux_radius(nu, 0.5), rel=1e-12)
    assert ksq_max == pytest.approx(p.ksq_max, rel=1e-14)
    s = g.Spergel(nu, half_light_radius=1.7, flux=2.5)
    assert s.scale_radius == pytest.approx(1.7 / pr.spergel_flux_radius(nu, 0.5), rel=1e-12)
    assert s.calculateIntegratedFlux(s.half_light_radius) == pytest.approx(0.5, abs=1e-13)


@pytest.mark.parametrize("nu", [-0.6, 0.5, 2.0])
@pytest.mark.parametrize("jac", [None, np.diag([1.3, 0.7]),
                                 np.array([[1.1, 0.3], [-0.2, 0.9]]) * 20.])
def test_clip_rows_match_oracle_fill(nu, jac):
    Nk, dk, r0 = 128, 2 * np.pi / 128, {-0.6: 0.05, 0.5: 15., 2.0: 2.}[nu]
    prof = pr.Spergel(nu, scale_radius=r0)
    out = np.zeros(2 * (Nk + 1), dtype=np.int32)
    j = None if jac is None else np.ascontiguousarray(jac.reshape(4))
    _lib.check(_lib.lib().gsmi_kclip_rows(prof.ksq_max, r0, _lib.ptr(j), Nk, dk, _lib.ptr(out)))
    img = od.KImage(0, Nk // 2, -Nk // 2, Nk // 2, dk)
    od.draw_k(prof, img, jac)
    mask = img.array != 0
    rows = out.reshape(-1, 2)
    if nu > 0:
        assert (~mask).any()            # the clipping radius lies inside the grid
    for jrow in range(Nk + 1):
        i1, i2 = rows[jrow]
        expect = np.zeros(Nk // 2 + 1, bool)
        expect[i1:i2] = True
        np.testing.assert_array_equal(mask[jrow], expect)


def test_range_and_argument_errors():
    with pytest.raises(ValueError):
        g.Spergel(-0.9, half_light_radius=1.)
    with pytest.raises(ValueError):
        g.Spergel(4.1, half_light_radius=1.)
    with pytest.raises(ValueError):
        g.Spergel(0., scale_radius=3., half_light_radius=1.)
    with pytest.raises(ValueError):
        g.Spergel(0.)
    rc = _lib.lib().gsmi_spergel_info(5.0, None, _lib.ptr(np.zeros(4)))
    assert rc == -2
