"""Real-space drawing (drawImage method='sb' / 'no_pixel' of analytic profiles) on the CPU: the
oracle pinned against the reference's saved images and closed forms, and galsim_amd's host-side
component records (galsim_amd.real.components) checked against the oracle through a numpy
evaluation of the records (no GPU)."""
import math
import os

import numpy as np
import pytest

import oracle.draw as od
import oracle.xdraw as ox
from tests.fitsio import read_fits

golden = os.path.join(os.path.dirname(__file__), "golden", "sb")
golden_top = os.path.join(os.path.dirname(__file__), "golden")


def _centred(N):
    """Bounds of an N x N image after setCenter(0, 0) (the centre sits right of / above the
    true centre for even N)."""
    return -(N // 2), N - 1 - N // 2, -(N // 2), N - 1 - N // 2


def _oracle_sb(obj, N, scale=0.2, method='sb'):
    return ox.draw_image_sb(obj, *_centred(N), scale, method=method)


# The reference's known-result tests (tests/test_exponential.py:29-45, test_moffat.py:30-48,
# test_sersic.py:30-95, test_box.py:28-40, 140-155, test_airy.py:29-45, test_spergel.py:29-55,
# test_kolmogorov.py:29-52,
# test_gaussian.py:30-40, test_transforms.py:45-383, test_sum.py:28-45): method='sb',
# use_true_center=False, scale 0.2, 5 decimals.
def _golden_cases():
    r0 = 1. / 1.67839
    fwhm_bc = 1.3178976627539716
    return [
        ("gauss_1.fits", lambda: od.Gaussian(flux=1., sigma=1.), golden_top),
        ("exp_1.fits", lambda: od.Exponential(flux=1., scale_radius=r0), golden),
        ("moffat_2_5.fits", lambda: od.Moffat(2., half_light_radius=1., trunc=5 * fwhm_bc,
                                              flux=1.), golden),
        ("sersic_3_1.fits", lambda: od.Sersic(3., flux=1., half_light_radius=1.), golden),
        ("sersic_3_1_10.fits", lambda: od.Sersic(3., flux=1., half_light_radius=1.,
                                                 trunc=9.99), golden),
        ("box_1.fits", lambda: od.Pixel(1., flux=1.), golden),
        ("tophat_101.fits", lambda: od.TopHat(1.01, flux=1.), golden),
        ("gauss_smallshear.fits", lambda: od.Gaussian(flux=1., sigma=1.).shear(e1=0.02,
                                                                                e2=0.02), golden),
        ("sersic_largeshear.fits", lambda: od.Sersic(4., flux=1., half_light_radius=1.).shear(
            e1=0., e2=0.5), golden),
        ("sersic_ellip_rotated.fits", lambda: od.Sersic(2.5, flux=1., half_light_radius=1.)
         .shear(e1=0.2, e2=0.).rotate(math.radians(45.)), golden),
        ("exp_mag.fits", lambda: od.Exponential(flux=1., scale_radius=r0).dilate(1.5)
         * 1.5 ** 2, golden),
        ("box_shift.fits", lambda: od.Pixel(0.2).shift(0.2, -0.2), golden),
        ("sersic_doubleflux.fits", lambda: od.Sersic(3., flux=2., half_light_radius=1.),
         golden),
        ("double_gaussian.fits", lambda: od.Add(od.Gaussian(flux=0.75, sigma=1.),
                                                od.Gaussian(flux=0.25, sigma=3.)), golden),
        ("airy_.8_.1.fits", lambda: od.Airy(1. / 0.8, obscuration=0.1, flux=1.), golden),
        ("kolmogorov.fits", lambda: od.Kolmogorov(lam_over_r0=1.5, flux=1.8), golden_top),
    ] + [("spergel_nu%.2f.fits" % nu, (lambda nu=nu: od.Spergel(nu, half_light_radius=1.)),
          golden) for nu in (-0.85, -0.5, 0.0, 0.85, 4.0)]


@pytest.mark.parametrize("name,make,where", _golden_cases(), ids=lambda v: v
                         if isinstance(v, str) and v.endswith('.fits') else '')
def test_oracle_matches_reference_saved_images(name, make, where):
    saved = read_fits(os.path.join(where, name))
    img = _oracle_sb(make(), saved.shape[0])
    np.testing.assert_array_almost_equal(img, saved, 5, err_msg=name)


def test_oracle_closed_forms():
    # peak of a centred Gaussian in surface brightness: flux / (2 pi sigma^2)
    img = _oracle_sb(od.Gaussian(flux=3., sigma=1.3), 64, scale=0.25)
    assert abs(img.max() - 3. / (2. * math.pi * 1.3 ** 2)) < 1e-15 * img.max() + 1e-16
    # flux of a well-sampled no_pixel drawing
    img = _oracle_sb(od.Exponential(flux=2.5, scale_radius=0.7), 256, scale=0.1,
                     method='no_pixel')
    assert abs(img.sum() - 2.5) < 2e-3
    # a sheared, rotated, shifted Gaussian against its closed form at every pixel centre
    sig, g1, g2, th, sx, sy = 0.9, 0.23, -0.11, 0.4, 0.31, -0.17
    obj = od.Gaussian(flux=1.7, sigma=sig).shear(g1=g1, g2=g2).rotate(th).shift(sx, sy)
    xmin, xmax, ymin, ymax = _centred(40)
    scale = 0.2
    img = ox.draw_image_sb(obj, xmin, xmax, ymin, ymax, scale)
    S = od.shear_matrix(g1=g1, g2=g2)
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    J = R @ S
    Ji = np.linalg.inv(J)
    x = (np.arange(xmin, xmax + 1) * scale)[None, :] - sx
    y = (np.arange(ymin, ymax + 1) * scale)[:, None] - sy
    u = Ji[0, 0] * x + Ji[0, 1] * y
    v = Ji[1, 0] * x + Ji[1, 1] * y
    want = 1.7 / (2. * math.pi * sig ** 2) / abs(np.linalg.det(J)) * np.exp(
        -0.5 * (u * u + v * v) / sig ** 2)
    np.testing.assert_allclose(img, want, rtol=0, atol=1e-13 * want.max())
    # Moffat with beta = 3 takes 1 / x^3; a beta just off 3 the general power
    for beta in (3., 3.2):
        m = od.Moffat(beta, scale_radius=0.8, flux=1.)
        img = _oracle_sb(m, 32, scale=0.2)
        x = (np.arange(-16, 16) * 0.2)
        rsq = (x[None, :] ** 2 + x[:, None] ** 2) / 0.64
        norm = (beta - 1.) / (math.pi * 0.64)
        np.testing.assert_allclose(img, norm * (1. + rsq) ** -beta, rtol=1e-13)


def _eval_records(recs, m, n, tables=()):
    """numpy evaluation of gsmi_xcomp records (the device formula restated for the host check)."""
    from galsim_amd import _lib
    ii = np.arange(m)[None, :].astype(float)
    jj = np.arange(n)[:, None].astype(float)
    total = None
    for c in recs:
        if c['form'] == 0:
            x = (np.abs(ii - c['izero']) * c['dx'] if c['izero'] else c['x0'] + ii * c['dx'])
            y = (np.abs(jj - c['jzero']) * c['dy'] if c['jzero'] else c['y0'] + jj * c['dy'])
            x, y = np.broadcast_arrays(x, y)
        else:
            x = (c['x0'] + jj * c['dxy']) + ii * c['dx']
            y = (c['y0'] + jj * c['dy']) + ii * c['dyx']
        t = int(c['type'])
        with np.errstate(divide='ignore'):
            if t == _lib.LEAF_GAUSSIAN:
                v = c['norm'] * np.exp(-0.5 * (x * x + y * y))
            elif t == _lib.LEAF_EXPONENTIAL:
                v = c['norm'] * np.exp(-np.sqrt(x * x + y * y))
            elif t == _lib.LEAF_MOFFAT:
                rsq = x * x + y * y
                v = np.where(rsq <= c['p'][1], c['norm'] * (1. + rsq) ** -c['p'][0], 0.)
            elif t == _lib.LEAF_SERSIC:
                rsq = x * x + y * y
                if c['ic'] >= 0:
                    rsq = rsq.copy()
                    rsq[c['jc'], c['ic']] = 0.
                v = c['norm'] * np.exp(-np.exp(c['p'][0] * np.log(rsq)))
                if c['p'][1] > 0:
                    v = np.where(rsq > c['p'][1], 0., v)
            elif t == _lib.LEAF_KOLMOGOROV:
                from oracle.numerics import SplineTable
                tab = np.concatenate(tables)[c['i1']:c['i1'] + 3 * c['i2']]
                k = int(c['i2'])
                sp = SplineTable(tab[:k], tab[k:2 * k])
                r = np.sqrt(x * x + y * y)
                v = c['norm'] * np.array([sp.lookup(float(q)) if q < tab[k - 1] else 0.
                                          for q in r.ravel()]).reshape(r.shape)
            elif t == _lib.LEAF_SPERGEL:
                import scipy.special as sps
                r = np.sqrt(x * x + y * y)
                nu = c['p'][0]
                v = c['norm'] * np.where(r == 0., c['p'][1], sps.kv(nu, r) * np.exp(nu * np.log(r)))
            elif t == _lib.LEAF_AIRY:
                import scipy.special as sps
                e = c['p'][0]
                nu = np.sqrt(x * x + y * y) * math.pi
                xv = np.where(nu < c['p'][1], 0.5 * (1. - e * e),
                              (sps.j1(nu) - e * sps.j1(e * nu)) / np.where(nu > 0, nu, 1.))
                v = c['norm'] * ((xv * xv) * (math.pi / (1. - e * e)))
            elif t == _lib.LEAF_BOX:
                if c['form'] == 0:
                    inside = ((ii >= c['i1']) & (ii < c['i2']) & (jj >= c['j1']) & (jj < c['j2']))
                else:
                    inside = (np.abs(x) < c['p'][0]) & (np.abs(y) < c['p'][1])
                v = np.where(inside, c['norm'], 0.)
            elif t == _lib.LEAF_TOPHAT:
                if c['form'] == 0:
                    yy = c['y0'] + jj * c['dy']
                    xmax = np.sqrt(np.maximum(c['p'][0] - yy * yy, 0.))
                    a = np.maximum(0, np.ceil(-xmax / abs(c['dx']) - c['x0'] / c['dx']))
                    b = np.minimum(m, np.ceil(xmax / abs(c['dx']) - c['x0'] / c['dx']))
                    inside = ((ii >= a) & (ii < b) & (jj >= c['j1']) & (jj < c['j2'])
                              & (c['p'][0] - yy * yy >= 0))
                else:
                    inside = x * x + y * y < c['p'][0]
                v = np.where(inside, c['norm'], 0.)
        v = np.broadcast_to(v, (n, m)) * c['amp']
        total = v if total is None else total + v
    return total


def _host_cases():
    import galsim_amd as g
    return [
        ("gauss", lambda m: m.Gaussian(flux=1.3, sigma=0.8)),
        ("exp_shift", lambda m: m.Exponential(flux=2., scale_radius=0.6).shift(0.13, -0.07)),
        ("moffat_trunc_shear", lambda m: m.Moffat(2.5, scale_radius=0.9, trunc=3., flux=1.)
         .shear(g1=0.2, g2=0.1)),
        ("moffat_untrunc_diag", lambda m: m.Moffat(3.7, half_light_radius=0.9).dilate(1.3)),
        ("sersic_rot", lambda m: m.Sersic(2.5, half_light_radius=1.).shear(g1=0.1, g2=0.)
         .rotate(0.6)),
        ("sersic_trunc", lambda m: m.Sersic(1.5, half_light_radius=0.8, trunc=3.)),
        ("box_shift", lambda m: m.Box(0.9, 0.5).shift(0.05, 0.03)),
        ("box_rot", lambda m: m.Box(1.1, 0.7).rotate(0.3)),
        ("tophat", lambda m: m.TopHat(0.93, flux=2.)),
        ("airy_obs_shear", lambda m: m.Airy(lam_over_diam=0.45, obscuration=0.3, flux=2.)
         .shear(g1=0.1, g2=0.05)),
        ("airy_unobs", lambda m: m.Airy(lam_over_diam=0.6)),
        ("spergel_shear", lambda m: m.Spergel(0.6, half_light_radius=0.9, flux=1.5)
         .shear(g1=0.2, g2=-0.1)),
        ("spergel_neg", lambda m: m.Spergel(-0.5, scale_radius=0.4)),
        ("kolm_shear", lambda m: m.Kolmogorov(fwhm=0.7, flux=1.3).shear(g1=0.1, g2=0.2)),
        ("kolm_sum", lambda m: m.Add(m.Kolmogorov(lam_over_r0=0.5),
                                     m.Gaussian(sigma=0.3, flux=0.2))),
        ("tophat_shear", lambda m: m.TopHat(0.93).shear(g1=-0.2, g2=0.15)),
        ("sum", lambda m: m.Add(m.Gaussian(sigma=0.5, flux=0.4),
                                m.Exponential(half_light_radius=1.1, flux=0.6).shear(g1=0.3,
                                                                                    g2=0.))),
    ]


@pytest.mark.parametrize("name,make", _host_cases(), ids=[c[0] for c in _host_cases()])
def test_host_records_match_oracle(name, make):
    import galsim_amd as g
    from galsim_amd import real
    from galsim_amd.position import BoundsI
    from galsim_amd.transform import _Transform
    scale, N = 0.2, 36
    obj = make(g)
    prof = _Transform(obj, np.array([[1. / scale, 0.], [0., 1. / scale]]), (0.21, -0.37),
                      scale ** 2 / scale ** 2)
    xmin, xmax, ymin, ymax = _centred(N)
    tables = []
    recs = real.components(prof, BoundsI(xmin, xmax, ymin, ymax), 1., tables)
    got = _eval_records(recs, N, N, tables)
    want = ox.draw_image_sb(make(od), xmin, xmax, ymin, ymax, scale, offset=(0.21, -0.37))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_records_reject_unsupported_profiles():
    import galsim_amd as g
    from galsim_amd import real
    from galsim_amd.position import BoundsI
    with pytest.raises(NotImplementedError):
        real.components(g.InclinedExponential(0.3 * g.radians, scale_radius=1., scale_height=0.1),
                        BoundsI(-4, 3, -4, 3))
    with pytest.raises(NotImplementedError):
        real.components(g.Convolve(g.Gaussian(sigma=1.), g.Box(1., 1.), real_space=True),
                        BoundsI(-4, 3, -4, 3))


def test_kolmogorov_xtable_matches_oracle():
    # the host table (libgsmi) against the oracle's restatement of the same loop
    import galsim_amd as g
    from galsim_amd import real
    import oracle.profiles as pr
    t = real.kolmogorov_xtable(g.GSParams())
    args, vals = pr.kolmogorov_xtable(pr.DEFAULT_GSPARAMS)
    n = len(t) // 3
    assert n == len(args)
    np.testing.assert_array_equal(t[:n], args)
    np.testing.assert_allclose(t[n:2 * n], vals, rtol=0, atol=1e-13 * vals[0])


# -- sheared Box / TopHat rows whose first pixel in the closed region lies on the edge ----------
# The reference fills such a row as zeros while outside, the norm while strictly inside, zeros
# after (src/SBBox.cpp:136-143, 358-362); a row whose first non-outside pixel is exactly on the
# edge stays zero.  Pixel(1) with x and y swapped, drawn at scale 0.5, puts the box edges on pixel
# centres: every row's first non-outside pixel is an edge pixel, so the reference draws zeros.
def test_box_edge_rows_stay_zero_in_oracle():
    obj = od.Pixel(1.).transform(0., 1., 1., 0.)
    img = ox.draw_image_sb(obj, *_centred(8), 0.5, method='no_pixel')
    assert np.all(img == 0.)
    # a per-pixel inside test would fill the centre pixel: the row rule is what empties it
    assert ox.draw_image_sb(od.Pixel(1.).transform(0., 1., 1., 0.), *_centred(8), 0.499,
                            method='no_pixel').max() > 0.


def _row_rule(cls):
    """The GPU kernel's per-pixel form of the row scan (gsmi_real.hip edge_row_filled): pixel i
    is filled iff it is strictly inside (class 2) and the pixel before its strictly-inside run
    (found by bisection) is outside (class 0), or the run starts the row."""
    out = np.zeros(len(cls), dtype=bool)
    for i in range(len(cls)):
        if cls[i] != 2:
            continue
        lo, hi = 0, i
        while lo < hi:
            mid = (lo + hi) // 2
            if cls[mid] == 2:
                hi = mid
            else:
                lo = mid + 1
        out[i] = lo == 0 or cls[lo - 1] == 0
    return out


@pytest.mark.parametrize("kind", ["box", "tophat"])
def test_kernel_row_rule_matches_reference_scan(kind):
    rng = np.random.default_rng(7)
    m = n = 24
    ncheck = 0
    for trial in range(60):
        # dyadic steps and origins, so that edges fall exactly on pixel centres in many rows
        d = rng.choice([0.25, 0.5, 0.125], 4) * rng.choice([-1., 1.], 4)
        if abs(d[0] * d[3] - d[1] * d[2]) < 1e-3:
            continue
        x0, y0 = -m / 2 * d[0] - n / 2 * d[1], -m / 2 * d[2] - n / 2 * d[3]
        X, Y = ox._grid(x0, d[0], d[1], y0, d[3], d[2], m, n)
        if kind == 'box':
            w, h = rng.choice([0.5, 1., 1.5, 0.75], 2)
            outside = (np.abs(X) > w) | (np.abs(Y) > h)
            inside = (np.abs(X) < w) & (np.abs(Y) < h)
        else:
            r2 = rng.choice([1., 2.25, 0.5625])
            outside = X * X + Y * Y > r2
            inside = X * X + Y * Y < r2
        cls = np.where(outside, 0, np.where(inside, 2, 1))
        for j in range(n):
            ref = np.zeros(m, dtype=bool)
            i = 0
            while i < m and outside[j, i]:
                i += 1
            while i < m and inside[j, i]:
                ref[i] = True
                i += 1
            assert np.array_equal(_row_rule(cls[j]), ref), (trial, j)
            ncheck += int(np.any(cls[j] == 1))
    assert ncheck > 50      # rows with exact-edge pixels were exercised
