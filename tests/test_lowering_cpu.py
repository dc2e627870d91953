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
