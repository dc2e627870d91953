"""Real-space drawing on the GPU (gsmi_draw_real_f64 through drawImage method='sb' / 'no_pixel'
and GSObject.drawReal) against the oracle, FP64, within 1e-10 of the peak pixel, and against the
reference's saved images at the reference's 5 decimals."""
import math
import os

import numpy as np
import pytest

import galsim_amd as g
import oracle.draw as od
import oracle.xdraw as ox
from galsim_amd.position import BoundsI
from tests.fitsio import read_fits
from tests.test_real_cpu import _centred, _golden_cases, _host_cases

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10


def _peak_rel(a, b):
    """max |a - b| over the largest |b|, over the finite pixels of b (a Spergel with nu <= 0 is
    infinite at r = 0 in both); the infinite pixels must agree exactly."""
    a = np.asarray(a, dtype=np.float64)
    fin = np.isfinite(b)
    assert np.array_equal(a[~fin], b[~fin])
    return float(np.max(np.abs(a[fin] - b[fin])) / np.max(np.abs(b[fin])))


def _golden_for_product(name):
    r0 = 1. / 1.67839
    fwhm_bc = 1.3178976627539716
    return {
        "gauss_1.fits": lambda: g.Gaussian(flux=1., sigma=1.),
        "exp_1.fits": lambda: g.Exponential(flux=1., scale_radius=r0),
        "moffat_2_5.fits": lambda: g.Moffat(beta=2, half_light_radius=1, trunc=5 * fwhm_bc,
                                            flux=1),
        "sersic_3_1.fits": lambda: g.Sersic(n=3, flux=1, half_light_radius=1),
        "sersic_3_1_10.fits": lambda: g.Sersic(n=3, flux=1, half_light_radius=1, trunc=9.99),
        "box_1.fits": lambda: g.Pixel(scale=1, flux=1),
        "tophat_101.fits": lambda: g.TopHat(radius=1.01, flux=1),
        "gauss_smallshear.fits": lambda: g.Gaussian(flux=1, sigma=1).shear(e1=0.02, e2=0.02),
        "sersic_largeshear.fits": lambda: g.DeVaucouleurs(flux=1, half_light_radius=1).shear(
            e1=0., e2=0.5),
        "sersic_ellip_rotated.fits": lambda: g.Sersic(n=2.5, flux=1, half_light_radius=1)
        .shear(e1=0.2, e2=0.).rotate(45. * g.degrees),
        "exp_mag.fits": lambda: g.Exponential(flux=1, scale_radius=r0).dilate(1.5) * 1.5 ** 2,
        "box_shift.fits": lambda: g.Pixel(scale=0.2).shift(0.2, -0.2),
        "sersic_doubleflux.fits": lambda: g.Sersic(n=3, flux=1, half_light_radius=1)
        .withFlux(2),
        "double_gaussian.fits": lambda: g.Add(g.Gaussian(flux=0.75, sigma=1),
                                              g.Gaussian(flux=0.25, sigma=3)),
        "airy_.8_.1.fits": lambda: g.Airy(lam_over_diam=1. / 0.8, obscuration=0.1, flux=1),
        "kolmogorov.fits": lambda: g.Kolmogorov(lam_over_r0=1.5, flux=1.8),
        "spergel_nu-0.85.fits": lambda: g.Spergel(nu=-0.85, half_light_radius=1.0),
        "spergel_nu-0.50.fits": lambda: g.Spergel(nu=-0.5, half_light_radius=1.0),
        "spergel_nu0.00.fits": lambda: g.Spergel(nu=0.0, half_light_radius=1.0),
        "spergel_nu0.85.fits": lambda: g.Spergel(nu=0.85, half_light_radius=1.0),
        "spergel_nu4.00.fits": lambda: g.Spergel(nu=4.0, half_light_radius=1.0),
    }[name]


@pytest.mark.parametrize("name,make,where", _golden_cases(),
                         ids=[c[0] for c in _golden_cases()])
def test_reference_saved_images(name, make, where):
    saved = read_fits(os.path.join(where, name))
    img = g.ImageF(saved.shape[1], saved.shape[0], scale=0.2)
    img.setCenter(0, 0)
    _golden_for_product(name)().drawImage(img, scale=0.2, method="sb", use_true_center=False)
    np.testing.assert_array_almost_equal(img.array, saved, 5, err_msg=name)
    # and the FP64 drawing against the oracle's
    imd = g.ImageD(saved.shape[1], saved.shape[0], scale=0.2)
    imd.setCenter(0, 0)
    _golden_for_product(name)().drawImage(imd, scale=0.2, method="sb", use_true_center=False)
    ref = ox.draw_image_sb(make(), *_centred(saved.shape[0]), 0.2)
    err = _peak_rel(imd.array, ref)
    print(name, "peak-rel", err)
    assert err < F64_TOL


@pytest.mark.parametrize("name,make", _host_cases(), ids=[c[0] for c in _host_cases()])
@pytest.mark.parametrize("method", ["no_pixel", "sb"])
@pytest.mark.parametrize("shape", [(36, 36), (41, 29)])
def test_drawimage_matches_oracle(name, make, method, shape):
    nx, ny = shape
    scale = 0.2
    offset = (0.21, -0.37)
    im = make(g).drawImage(nx=nx, ny=ny, scale=scale, dtype=np.float64, method=method,
                           offset=offset)
    # drawImage's own offset handling: use_true_center shifts even sizes by half a pixel
    ox_, oy_ = offset[0] - (0.5 if nx % 2 == 0 else 0.), offset[1] - (0.5 if ny % 2 == 0 else 0.)
    b = im.bounds
    c = im.center
    ref = ox.draw_image_sb(make(od), b.xmin - c.x, b.xmax - c.x, b.ymin - c.y, b.ymax - c.y,
                           scale, method=method, offset=(ox_, oy_))
    err = _peak_rel(im.array, ref)
    print(name, method, shape, "peak-rel", err)
    assert err < F64_TOL
    want = float(ref.sum()) / ((1. / scale ** 2) if method == 'sb' else 1.)
    assert im.added_flux == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_drawreal_direct_scale_and_add():
    obj = g.Exponential(half_light_radius=0.8, flux=3.).shear(g1=0.15, g2=0.05)
    img = g.ImageD(bounds=BoundsI(-20, 19, -15, 16), scale=0.25)
    added = obj.drawReal(img)
    ref = ox.draw_real(od.Exponential(half_light_radius=0.8, flux=3.).shear(g1=0.15, g2=0.05),
                       -20, 19, -15, 16, 0.25)
    assert _peak_rel(img.array, ref) < F64_TOL
    assert added == pytest.approx(float(ref.sum()), rel=1e-12)
    # adding goes through a float32 temporary, as in the reference (galsim/gsobject.py:1879-1888)
    added2 = obj.drawReal(img, add_to_image=True)
    ref32 = ref.astype(np.float32)
    assert _peak_rel(img.array, ref + ref32) < F64_TOL
    assert added2 == pytest.approx(float(ref32.sum(dtype=float)), rel=1e-12)


def test_sum_of_many_components_chunks():
    # more components than one launch takes (16): drawn in chunks that accumulate in order
    parts_g = [g.Gaussian(sigma=0.3 + 0.05 * k, flux=1. + 0.1 * k).shift(0.05 * k, -0.03 * k)
               for k in range(21)]
    parts_o = [od.Gaussian(sigma=0.3 + 0.05 * k, flux=1. + 0.1 * k).shift(0.05 * k, -0.03 * k)
               for k in range(21)]
    im = g.Add(parts_g).drawImage(nx=48, ny=48, scale=0.1, dtype=np.float64, method='no_pixel',
                                  use_true_center=False)
    ref = ox.draw_image_sb(od.Add(*parts_o), *_centred(48), 0.1, method='no_pixel')
    assert _peak_rel(im.array, ref) < F64_TOL


def test_large_image_flux():
    # a size-independent property at a large size: the no_pixel flux of a well-sampled profile
    obj = g.Gaussian(sigma=20., flux=7.).shear(g1=0.2, g2=-0.1)
    im = obj.drawImage(nx=2048, ny=2048, scale=0.2, dtype=np.float64, method='no_pixel')
    assert float(im.array.sum()) == pytest.approx(7., rel=1e-9)
    assert im.added_flux == pytest.approx(7., rel=1e-9)


def test_float32_image():
    obj = g.Sersic(n=4, half_light_radius=0.7).shear(g1=0.1, g2=0.2)
    im = obj.drawImage(nx=40, ny=40, scale=0.2, method='sb')
    assert im.array.dtype == np.float32
    ref = ox.draw_image_sb(od.Sersic(4., half_light_radius=0.7).shear(g1=0.1, g2=0.2),
                           *_centred(40), 0.2, offset=(-0.5, -0.5))
    assert _peak_rel(im.array, ref) < 3e-7


def test_sheared_box_edge_rows_follow_reference_scan():
    # the box edges fall on pixel centres: every row's first non-outside pixel is an edge pixel,
    # so the reference's row scan (src/SBBox.cpp:136-143) leaves the image zero
    im = g.Pixel(scale=1.).transform(0., 1., 1., 0.).drawImage(
        nx=8, ny=8, scale=0.5, method='no_pixel', dtype=np.float64, use_true_center=False)
    assert np.all(im.array == 0.)
    for name, obj_g, obj_o in [
            ("box", g.Box(1.5, 1.).transform(0.5, 0.25, -0.25, 0.5),
             od.Box(1.5, 1.).transform(0.5, 0.25, -0.25, 0.5)),
            ("tophat", g.TopHat(1.5).transform(0.5, 0.25, 0.25, 1.),
             od.TopHat(1.5).transform(0.5, 0.25, 0.25, 1.))]:
        im = obj_g.drawImage(nx=24, ny=24, scale=0.25, method='no_pixel', dtype=np.float64,
                             use_true_center=False)
        ref = ox.draw_image_sb(obj_o, *_centred(24), 0.25, method='no_pixel')
        assert np.array_equal(im.array != 0., ref != 0.), name
        assert _peak_rel(im.array, ref) < F64_TOL, name
