"""Config-driven batches (galsim_amd.config): the vectorised planner gives the records of the
per-object planner (galsim_amd.batch.plan_host) on the profiles the same config values build,
config values evaluate as GalSim's config defines them, and unsupported fields raise."""
import math

import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import _lib, batch, config


def _compare(hp, ho, stride=1):
    a, b = hp.recs, ho.recs
    assert np.array_equal(hp.keep, ho.keep)
    keep = np.nonzero(hp.keep)[0]
    for f in ('N', 'Nk', 'nx', 'ny', 'x0', 'y0', 'op_count', 'axis_count'):
        assert np.array_equal(a[f][keep], b[f][keep]), f
    np.testing.assert_allclose(a['dk'][keep], b['dk'][keep], rtol=1e-15, atol=0)
    worst = 0.
    for i in keep[::stride]:
        na = int(a['op_count'][i] + a['axis_count'][i])
        oa = hp.ops[a['op_begin'][i]:a['op_begin'][i] + na]
        ob = ho.ops[b['op_begin'][i]:b['op_begin'][i] + na]
        assert np.array_equal(oa['code'], ob['code']) and np.array_equal(oa['arg'], ob['arg'])
        for f in 'abc':
            worst = max(worst, float(np.max(np.abs(oa[f] - ob[f]) /
                                            np.maximum(np.abs(ob[f]), 1e-300))))
        nl = int(oa['arg'][oa['code'] == _lib.OP_LEAF].max()) + 1
        la = hp.leaves[a['leaf_begin'][i]:a['leaf_begin'][i] + nl]
        lb = ho.leaves[b['leaf_begin'][i]:b['leaf_begin'][i] + nl]
        for f in ('type', 'separable', 'kind', 'ntab'):
            assert np.array_equal(la[f], lb[f]), (i, f)
        for f in ('scale', 'scale2', 'flux', 'p'):
            worst = max(worst, float(np.max(np.abs(la[f] - lb[f]) /
                                            np.maximum(np.abs(lb[f]), 1e-300))))
        # Jacobian entries relative to the matrix (an entry can be a near-cancellation of
        # products, where numpy's and the per-object matrix product round differently)
        scale = np.max(np.abs(lb['jac']), axis=1, keepdims=True)
        worst = max(worst, float(np.max(np.abs(la['jac'] - lb['jac']) / scale)))
        # tables, FP32 spline coefficients and FP64 clip rows read the same values
        for k in range(nl):
            if la['type'][k] == _lib.LEAF_SERSIC:
                n3 = 3 * int(la['ntab'][k])
                assert np.array_equal(hp.tables[la['table'][k]:la['table'][k] + n3],
                                      ho.tables[lb['table'][k]:lb['table'][k] + n3])
                if la['cliprows'][k] >= 0:
                    m = 4 * (int(la['ntab'][k]) - 1)
                    assert np.array_equal(hp.tables32[la['cliprows'][k]:la['cliprows'][k] + m],
                                          ho.tables32[lb['cliprows'][k]:lb['cliprows'][k] + m])
            if la['type'][k] == _lib.LEAF_EXPONENTIAL and la['cliprows'][k] >= 0:
                m = 2 * (int(a['Nk'][i]) + 1)
                assert np.array_equal(hp.clip[la['cliprows'][k]:la['cliprows'][k] + m],
                                      ho.clip[lb['cliprows'][k]:lb['cliprows'][k] + m])
    # the same arithmetic up to the order numpy and Python take for a few products
    assert worst < 4e-16, worst


def bulge_disk_config(n, seed=1):
    rng = np.random.default_rng(seed)
    return {
        'gal': {'type': 'Sum', 'items': [
            {'type': 'Exponential', 'half_light_radius': np.exp(rng.normal(0., 0.4, n)),
             'flux': rng.uniform(100., 1e4, n)},
            {'type': 'Sersic', 'n': rng.choice([1.5, 2., 2.5, 3., 3.5, 4.], n),
             'half_light_radius': rng.uniform(0.3, 1.0, n), 'flux': rng.uniform(10., 1e3, n)}],
            'shear': {'type': 'G1G2', 'g1': rng.uniform(-0.3, 0.3, n),
                      'g2': rng.uniform(-0.3, 0.3, n)}},
        'psf': {'type': 'Kolmogorov', 'fwhm': {'type': 'Random', 'min': 0.6, 'max': 1.0}},
        'image': {'pixel_scale': 0.2},
        'stamp': {'size': {'type': 'List', 'items': [32, 48, 64, 96, 128]},
                  'offset': {'type': 'XY', 'x': {'type': 'Random', 'min': -0.5, 'max': 0.5},
                             'y': 0.1}}}


@pytest.mark.parametrize("precision", ['f32', 'f64'])
def test_bulge_disk_config_matches_per_object_planning(precision):
    n = 384
    cfg = bulge_disk_config(n)
    hp = config.plan_config(cfg, n, precision, seed=3)
    profs, nx, ny, scale, offs, method = config.build_profiles(cfg, n, seed=3)
    ho = batch.plan_host(profs, nx, ny, scale, offs, precision, method)
    _compare(hp, ho)


def test_transform_chain_moffat_and_values():
    # every transformation of TransformObject, a Moffat of varying beta (generic and special),
    # a flux on a Sum, Sequence / List / RandomGaussian / E1E2 / GBeta values
    n = 96
    rng = np.random.default_rng(5)
    beta = rng.uniform(2.5, 4.5, n)
    beta[::7] = 3.0                                       # the closed-form K_2 kind
    cfg = {
        'gal': {'type': 'Sum', 'flux': {'type': 'Sequence', 'first': 1000., 'step': 10.},
                'items': [{'type': 'Gaussian', 'sigma': 0.5, 'flux': 0.3},
                          {'type': 'Exponential', 'scale_radius': 0.6, 'flux': 0.7}],
                'dilate': {'type': 'RandomGaussian', 'mean': 1.0, 'sigma': 0.05},
                'ellip': {'type': 'E1E2', 'e1': 0.1, 'e2': {'type': 'Random', 'min': -0.2,
                                                              'max': 0.2}},
                'rotate': {'type': 'Random'},
                'scale_flux': 1.5,
                'shear': {'type': 'GBeta', 'g': 0.05, 'beta': 0.3},
                'magnify': 1.02,
                'shift': {'type': 'XY', 'x': 0.1, 'y': {'type': 'List',
                                                        'items': [0., -0.2, 0.3]}}},
        'psf': {'type': 'Moffat', 'beta': beta, 'fwhm': 0.7},
        'image': {'pixel_scale': 0.25},
        'stamp': {'xsize': 40, 'ysize': 48}}
    for precision in ('f32', 'f64'):
        hp = config.plan_config(cfg, n, precision, seed=9)
        profs, nx, ny, scale, offs, method = config.build_profiles(cfg, n, seed=9)
        ho = batch.plan_host(profs, nx, ny, scale, offs, precision, method)
        _compare(hp, ho)
    # values: a Sequence, a List indexed by the object number, the shift in the records
    V = config._Values(7, 0)
    assert list(V.num({'type': 'Sequence', 'first': 2, 'step': 3})) == [2, 5, 8, 11, 14, 17, 20]
    assert list(V.num({'type': 'List', 'items': [1, 2, 3]})) == [1, 2, 3, 1, 2, 3, 1]
    assert list(V.num({'type': 'Sequence', 'repeat': 2})) == [0, 0, 1, 1, 2, 2, 3]
    g1, g2 = V.shear({'type': 'GBeta', 'g': 0.2, 'beta': math.pi / 8})
    np.testing.assert_allclose(g1, 0.2 * math.cos(math.pi / 4))


def test_bad_values_are_flagged_and_unsupported_fields_raise():
    n = 5
    cfg = {'gal': {'type': 'Exponential', 'half_light_radius': [1., np.nan, 1.2, 0.9, 1.1],
                   'flux': [1., 1., np.inf, 1., 1.]},
           'psf': {'type': 'Kolmogorov', 'fwhm': 0.8},
           'image': {'pixel_scale': 0.2}, 'stamp': {'size': 32}}
    hp = config.plan_config(cfg, n, 'f32')
    assert list(hp.keep) == [True, False, False, True, True]
    assert hp.host_status[1] & batch.STAMP_BAD_PARAMS and hp.host_status[2] & batch.STAMP_BAD_PARAMS
    with pytest.raises(ValueError):
        config.plan_config({'gal': {'type': 'Sersic', 'n': 2, 'half_light_radius': 1.,
                                    'trunc': 3.}, 'stamp': {'size': 32}}, 2)
    with pytest.raises(ValueError):
        config.plan_config({'gal': {'type': 'Exponential', 'half_light_radius': 1.},
                            'stamp': {'size': 32, 'draw_method': 'phot'}}, 2)
    with pytest.raises(ValueError):
        config._Values(3, 0).num({'type': 'Catalog'})
