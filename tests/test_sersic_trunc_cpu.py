"""Truncated Sersic (src/SBSersic.cpp: SersicInfo with trunc > 0, buildFT's hankel_trunc branch,
CalculateTruncatedScale; galsim/sersic.py: flux_untruncated) on the host, against the
reference's own SersicInfo compiled from its sources (oracle/_ref), its known answers, and the
oracle's restatement."""
import math

import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import _lib
from oracle import profiles as op
from oracle import ref

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
# the reference integrates its truncated tables with adaptive GKP to relerr 1e-6: its k-values
# sit within ~3e-9 (of the k = 0 value) of the transform; ours and the oracle's within 1e-12
REF_TABLE_TOL = 5e-9


def _product(n, trunc):
    out = np.zeros(11)
    nt = np.zeros(1, dtype=np.int32)
    _lib.check(_lib.lib().gsmi_sersic_info(n, trunc, None, _lib.ptr(out), _lib.ptr(nt)))
    m = int(nt[0])
    a, v, y2 = np.zeros(m), np.zeros(m), np.zeros(m)
    _lib.check(_lib.lib().gsmi_sersic_table(n, trunc, None, _lib.ptr(a), _lib.ptr(v),
                                             _lib.ptr(y2), m))
    return out, a, v, y2


def _kvalues(n, trunc, ksq):
    out = np.zeros_like(ksq)
    _lib.check(_lib.lib().gsmi_sersic_kvalue(n, trunc, None, _lib.ptr(ksq), _lib.ptr(out),
                                              len(ksq)))
    return out


@needs_ref
@pytest.mark.parametrize("n,trunc", [(3.1, 100.87), (4.0, 20.), (1.5, 5.), (0.7, 3.),
                                     (6.0, 50.)])
def test_sizes_and_kvalues_match_reference(n, trunc):
    sc = _product(n, trunc)[0]
    maxk, stepk, hlr, _ = ref.sersic_info(n, trunc)
    assert sc[0] == pytest.approx(maxk, rel=1e-13)
    assert sc[1] == pytest.approx(stepk, rel=1e-13)
    assert sc[2] == pytest.approx(hlr, rel=1e-13)
    ksq = np.exp(np.linspace(np.log(1e-6), np.log(1e5), 2001))
    assert np.max(np.abs(_kvalues(n, trunc, ksq) - ref.sersic_kvalue(n, ksq, trunc))) \
        < REF_TABLE_TOL


@pytest.mark.parametrize("n,trunc", [(1.5, 5.), (4.0, 8.), (0.7, 3.), (3.1, 100.87)])
def test_tables_match_oracle(n, trunc):
    sc, a, v, y2 = _product(n, trunc)
    o = op.SersicInfo(n, op.DEFAULT_GSPARAMS, trunc)
    # the first knot is log(kmin) - 0.001 with kmin from the truncated gamma functions, whose
    # implementations (ours, scipy's) differ in the last bits
    np.testing.assert_allclose(a, o.table.args, rtol=0, atol=1e-13)
    assert np.max(np.abs(v - o.table.vals) / np.maximum(np.exp(2 * a), 1.)) < 1e-12
    assert sc[0] == pytest.approx(o.maxk, rel=1e-13) and sc[1] == pytest.approx(o.stepk)
    assert sc[10] == pytest.approx(o.flux_fraction, rel=1e-14)
    ksq = np.exp(np.linspace(np.log(1e-6), np.log(1e5), 2001))
    assert np.max(np.abs(_kvalues(n, trunc, ksq) - o.kvalue(ksq))) < 1e-12
    if ref.available():
        assert np.max(np.abs(o.kvalue(ksq) - ref.sersic_kvalue(n, ksq, trunc))) < REF_TABLE_TOL


def test_reference_known_answers():
    # tests/test_sersic.py:497-530: n = 3.1, hlr = 2.3, trunc = 2 hlr, flux 1.8
    p = g.Sersic(n=3.1, half_light_radius=2.3, trunc=4.6, flux=1.8)
    assert p.scale_radius == pytest.approx(0.04559626861897289, rel=1e-12)
    assert p.maxk == pytest.approx(24.663782559047256, rel=1e-12)
    assert p.stepk == pytest.approx(0.6829549246934334, rel=1e-12)
    assert p.half_light_radius == pytest.approx(2.3, rel=1e-12)
    q = g.Sersic(n=3.1, scale_radius=0.04559626861897289, trunc=4.6, flux=1.8)
    assert q.maxk == pytest.approx(24.663782559047256, rel=1e-12)
    assert q.half_light_radius == pytest.approx(2.3, abs=1e-7)
    # maxk and stepk scale with the radius
    q5 = g.Sersic(n=3.1, half_light_radius=5 * 2.3, trunc=5 * 4.6)
    assert q5.maxk == pytest.approx(p.maxk / 5, rel=1e-12)
    if ref.available():
        assert p.scale_radius == pytest.approx(ref.sersic_truncated_scale(3.1, 2.3, 4.6),
                                               rel=1e-14)


def test_flux_untruncated_and_argument_checks():
    # galsim/sersic.py:201-250: with flux_untruncated the given flux is the untruncated
    # profile's, so the drawn flux is flux * P(2n, (trunc/r0)^(1/n)), and r0 follows the
    # untruncated hlr relation
    p = g.Sersic(2.5, half_light_radius=1., trunc=3., flux=10., flux_untruncated=True)
    u = g.Sersic(2.5, half_light_radius=1., flux=10.)
    assert p.scale_radius == pytest.approx(u.scale_radius, rel=1e-15)
    ff = op._gamma_p(5., op.fast_pow(3. / p.scale_radius, 1. / 2.5))
    assert p.flux == pytest.approx(10. * ff, rel=1e-13)
    assert p.half_light_radius < 1.
    with pytest.raises(g.GalSimAMDRangeError):
        g.Sersic(2.5, half_light_radius=1., trunc=1.4)
    with pytest.raises(g.GalSimAMDRangeError):
        g.Sersic(2.5, half_light_radius=1., trunc=-1.)
