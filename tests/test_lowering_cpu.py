"""CPU check of the host planning and lowering: the lowered k-space program, evaluated by a
numpy emulation of the kernel's leaf formulas, reproduces the oracle's drawImage."""
import numpy as np
import pytest

import galsim_amd as g
from galsim_amd.position import BoundsI, PositionI
from galsim_amd.transform import _Transform
from oracle import draw as od
from tests.program_emulator import emulate_draw

CASES = [
    ("exp_moffat", lambda m: m.Convolve([m.Exponential(half_light_radius=1.5, flux=1e4)
                                         .shear(g1=0.1, g2=-0.05), m.Moffat(beta=3, fwhm=0.7)]),
     (64, 64), (0.13, -0.21)),
    ("gaussian_sheared", lambda m: m.Gaussian(sigma=0.7, flux=3.).shear(e1=0.3, e2=0.1),
     (40, 40), (0., 0.)),
    ("kolmogorov_folded", lambda m: m.Kolmogorov(fwhm=0.6, flux=2.), (32, 32), (0.2, 0.)),
    ("moffat_generic", lambda m: m.Moffat(2.7, fwhm=0.9).shear(g1=-0.2, g2=0.15), (48, 48),
     (0., 0.)),
    ("bulge_disk", lambda m: m.Convolve([
        (m.Exponential(half_light_radius=1.0, flux=600.) +
         m.Sersic(2.5, half_light_radius=0.5, flux=400.)).shear(g1=0.3, g2=-0.2).shift(0.1, 0.05),
        m.Kolmogorov(fwhm=0.75)]), (64, 64), (0.3, -0.4)),
    ("odd_nonsquare", lambda m: m.Convolve([m.Exponential(half_light_radius=1.1)
                                            .shear(g1=0.2, g2=0.3), m.Moffat(beta=3, fwhm=0.7)]),
     (37, 52), (0., 0.)),
    ("exp_small_clipped", lambda m: m.Exponential(scale_radius=0.05).shear(g1=0.4, g2=0.1),
     (24, 24), (0., 0.)),
]


@pytest.mark.parametrize("hoist", [False, True])
@pytest.mark.parametrize("name,make,shape,offset", CASES)
def test_lowered_program_matches_oracle(name, make, shape, offset, hoist):
    nx, ny = shape
    scale = 0.2
    prod = make(g)
    ref, info = make(od).drawImage(nx, ny, scale=scale, offset=offset, return_info=True)
    # the same steps as GSObject.drawImage, stopping before the device call
    im = prod.drawImage(nx=nx, ny=ny, scale=scale, offset=offset, setup_only=True)
    off = g.PositionD(offset)
    off = prod._adjust_offset(g.BoundsI(1, nx, 1, ny), off, None, True)
    jac = np.array([[1. / scale, 0.], [0., 1. / scale]])
    p = _Transform(prod, jac, (off.x, off.y), scale ** 2)
    p = g.Convolve(p, g.Pixel(scale=1.0, gsparams=prod.gsparams), real_space=False,
                   gsparams=prod.gsparams)
    c = im.center
    b = im.bounds.shift(PositionI(-c.x, -c.y))
    emu = emulate_draw(p, b, hoist=hoist)
    err = np.max(np.abs(emu - ref)) / np.max(np.abs(ref))
    assert err < 1e-10, (name, err)


# -- hard edges and drawImage(method='auto') (galsim/gsobject.py:1754-1764, convolve.py:150-175) --
def test_truncated_profiles_have_hard_edges():
    import galsim_amd as g
    assert g.Moffat(beta=2.5, scale_radius=1., trunc=4.).has_hard_edges
    assert not g.Moffat(beta=2.5, scale_radius=1.).has_hard_edges
    assert g.Sersic(n=2., half_light_radius=1., trunc=5.).has_hard_edges
    assert not g.Sersic(n=2., half_light_radius=1.).has_hard_edges
    assert g.Sersic(n=2., half_light_radius=1., trunc=5.).shear(g1=0.1, g2=0.).has_hard_edges


def test_auto_real_space_choice_for_hard_edges():
    import warnings
    import galsim_amd as g
    hard = [g.Box(1., 1.), g.Pixel(scale=1.), g.TopHat(1.),
            g.Moffat(beta=2.5, scale_radius=1., trunc=4.),
            g.Sersic(n=2., half_light_radius=1., trunc=5.)]
    for obj in hard:
        # two hard-edged profiles: the reference convolves in real space, which is not built here
        with pytest.raises(NotImplementedError):
            g.Convolve(obj, g.Pixel(scale=1.))
        with pytest.raises(NotImplementedError):
            g.AutoConvolve(obj)
        with pytest.raises(NotImplementedError):
            obj.drawImage(nx=16, ny=16, scale=0.2, setup_only=True)
        # method='fft' forces the FFT, with the reference's warning
        with pytest.warns(UserWarning, match="hard edges"):
            obj.drawImage(nx=16, ny=16, scale=0.2, method='fft', setup_only=True)
    soft = g.Moffat(beta=2.5, scale_radius=1.)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        soft.drawImage(nx=16, ny=16, scale=0.2, setup_only=True)
        g.Convolve(soft, g.Pixel(scale=1.))


def test_drawreal_temporary_dtype():
    from galsim_amd.gsobject import _real_values
    px = np.array([[1. + 2. ** -30, 2.], [3., 4.]])
    for dt, add, want in [(np.float64, False, np.float64), (np.float32, False, np.float32),
                          (np.float64, True, np.float32), (np.float32, True, np.float32),
                          (np.complex64, False, np.float32), (np.complex128, False, np.float64),
                          (np.int32, False, np.float64), (np.uint32, True, np.float64),
                          (np.int16, False, np.float32)]:
        vals, added = _real_values(px, np.dtype(dt), add)
        assert vals.dtype == want, (dt, add)
        assert added == float(px.astype(want).sum(dtype=float))


@pytest.mark.parametrize("idx", range(4))
def test_lowered_phases_match_closed_form(idx):
    # the product's lowering of the closed-form transformed Gaussians (both ApplyKImagePhases
    # branches, offsets composed through outer Jacobians), evaluated by the program emulator
    from galsim_amd import render
    from tests.program_emulator import eval_program
    from tests.test_oracle_phases import _cases, closed_form
    name, obj, terms = _cases(g)[idx]
    nx, ny, dk = 41, 37, 0.21
    ctx = render.Lowering('f64', dk=dk)
    obj._lower(ctx, None)
    ii, jj = np.meshgrid(np.arange(nx) - nx // 2, np.arange(ny) - ny // 2)
    got = eval_program(ctx, ii, jj, dk, 0)
    ref = closed_form(terms, nx, ny, dk)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    assert err < 1e-12, (name, err)
