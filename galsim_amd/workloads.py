"""Seeded synthetic stamp sets used by the benchmark and the batch tests.

Each builder takes a module ``m`` exposing the galsim-style constructors (``galsim_amd`` itself,
or any object with the same names) so that identical parameter draws can build matching profile
trees for different renderers.
"""
import math

import numpy as np


# -- Exponential (x) Moffat on 64x64 stamps --------------------------------------------------
def exp_moffat_params(n, seed=0, flux=1e4):
    """hlr ~ U(0.8, 2.5)", |g| ~ U(0, 0.5) at a uniform angle, sub-pixel offset ~ U(-0.5, 0.5)^2;
    stamp 0 is hlr = 1.5", (g1, g2) = (0.1, -0.05), offset (0.13, -0.21)."""
    rng = np.random.default_rng(seed)
    hlr = rng.uniform(0.8, 2.5, n)
    gmag = rng.uniform(0., 0.5, n)
    phi = rng.uniform(0., np.pi, n)
    g1 = gmag * np.cos(2 * phi)
    g2 = gmag * np.sin(2 * phi)
    off = rng.uniform(-0.5, 0.5, (n, 2))
    if n:
        hlr[0], g1[0], g2[0], off[0] = 1.5, 0.1, -0.05, (0.13, -0.21)
    return dict(hlr=hlr, g1=g1, g2=g2, offset=off, flux=np.full(n, flux), beta=np.full(n, 3.),
                psf_fwhm=np.full(n, 0.7))


def exp_moffat_profile(m, p, i):
    gal = m.Exponential(half_light_radius=float(p['hlr'][i]), flux=float(p['flux'][i]))
    gal = gal.shear(g1=float(p['g1'][i]), g2=float(p['g2'][i]))
    psf = m.Moffat(beta=float(p['beta'][i]), fwhm=float(p['psf_fwhm'][i]))
    return m.Convolve([gal, psf])


# -- bulge + disk (x) Kolmogorov, uneven stamp sizes --------------------------------------------
SERSIC_N = (1.5, 2.0, 2.5, 3.0, 3.5, 4.0)


def bulge_disk_params(n, seed=0):
    """Disk Exponential hlr ~ LogNormal(ln 1", 0.4); bulge Sersic n in {1.5, ..., 4.0} with
    hlr = 0.5 disk hlr and flux fraction ~ U(0, 0.6); |g| ~ U(0, 0.6) at a uniform angle;
    Kolmogorov fwhm ~ U(0.6, 1.0)"; total flux ~ LogUniform(1e2, 1e5)."""
    rng = np.random.default_rng(seed)
    disk_hlr = np.exp(rng.normal(np.log(1.0), 0.4, n))
    sn = rng.choice(np.array(SERSIC_N), n)
    bfrac = rng.uniform(0., 0.6, n)
    gmag = rng.uniform(0., 0.6, n)
    phi = rng.uniform(0., np.pi, n)
    fwhm = rng.uniform(0.6, 1.0, n)
    flux = np.exp(rng.uniform(np.log(1e2), np.log(1e5), n))
    return dict(disk_hlr=disk_hlr, n=sn, bfrac=bfrac, g1=gmag * np.cos(2 * phi),
                g2=gmag * np.sin(2 * phi), psf_fwhm=fwhm, flux=flux)


def bulge_disk_profile(m, p, i):
    F = float(p['flux'][i])
    f = float(p['bfrac'][i])
    dh = float(p['disk_hlr'][i])
    disk = m.Exponential(half_light_radius=dh, flux=(1. - f) * F)
    bulge = m.Sersic(float(p['n'][i]), half_light_radius=0.5 * dh, flux=f * F)
    gal = (disk + bulge).shear(g1=float(p['g1'][i]), g2=float(p['g2'][i]))
    return m.Convolve([gal, m.Kolmogorov(fwhm=float(p['psf_fwhm'][i]))])


def bulge_disk_config(p, sizes):
    """The same bulge + disk stamps as a config (galsim_amd.config): per-object values from the
    parameter arrays of bulge_disk_params, stamp sides ``sizes``."""
    F, f, dh = p['flux'], p['bfrac'], p['disk_hlr']
    return {'gal': {'type': 'Sum', 'items': [
                {'type': 'Exponential', 'half_light_radius': dh, 'flux': (1. - f) * F},
                {'type': 'Sersic', 'n': p['n'], 'half_light_radius': 0.5 * dh,
                 'flux': f * F}],
                    'shear': {'type': 'G1G2', 'g1': p['g1'], 'g2': p['g2']}},
            'psf': {'type': 'Kolmogorov', 'fwhm': p['psf_fwhm']},
            'image': {'pixel_scale': 0.2},
            'stamp': {'size': np.asarray(sizes, dtype=float)}}


# -- large Sersic stamps (x) Kolmogorov: FFT grids of 512..4096 --------------------------------
def large_sersic_params(n, seed=0):
    """Sersic n ~ U(3, 6) quantised to 0.05, hlr ~ U(2, 6)", Kolmogorov fwhm 0.8", flux 1e4."""
    rng = np.random.default_rng(seed)
    sn = np.round(rng.uniform(3.0, 6.0, n) / 0.05) * 0.05
    return dict(n=np.round(sn, 2), hlr=rng.uniform(2., 6., n), psf_fwhm=np.full(n, 0.8),
                flux=np.full(n, 1e4))


def large_sersic_profile(m, p, i):
    gal = m.Sersic(float(p['n'][i]), half_light_radius=float(p['hlr'][i]),
                   flux=float(p['flux'][i]))
    return m.Convolve([gal, m.Kolmogorov(fwhm=float(p['psf_fwhm'][i]))])


def stamp_size(prof, scale, lo=32, hi=256):
    """goodFFTSize(getGoodImageSize(scale)) clipped to [lo, hi]."""
    from .image import Image
    return int(min(max(Image.good_fft_size(prof.getGoodImageSize(scale)), lo), hi))


def build(kind, n, seed=0, m=None):
    """(profiles, nx, ny, offsets, params) for workload ``kind`` ('exp_moffat', 'bulge_disk' or
    'large_sersic')."""
    if m is None:
        import galsim_amd as m
    if kind == 'exp_moffat':
        p = exp_moffat_params(n, seed)
        profs = [exp_moffat_profile(m, p, i) for i in range(n)]
        nx = np.full(n, 64)
        return profs, nx, nx.copy(), p['offset'], p
    if kind == 'bulge_disk':
        p = bulge_disk_params(n, seed)
        profs = [bulge_disk_profile(m, p, i) for i in range(n)]
        import galsim_amd as g
        sizes = np.array([stamp_size(bulge_disk_profile(g, p, i), 0.2) for i in range(n)])
        return profs, sizes, sizes.copy(), np.zeros((n, 2)), p
    if kind == 'large_sersic':
        # drawImage's own size for each profile (getGoodImageSize at the pixel scale)
        p = large_sersic_params(n, seed)
        profs = [large_sersic_profile(m, p, i) for i in range(n)]
        import galsim_amd as g
        sizes = np.array([large_sersic_profile(g, p, i).getGoodImageSize(0.2)
                          for i in range(n)])
        return profs, sizes, sizes.copy(), np.zeros((n, 2)), p
    raise ValueError(kind)
