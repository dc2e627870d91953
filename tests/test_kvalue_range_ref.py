"""GetKValueRange1d / GetKValueRange2d (src/SBProfile.cpp:508-588) pinned bit for bit against the
reference's own compiled functions (oracle/_ref/libgsref.so): the oracle's restatement, and the
product's host clip rows (gsmi_exponential_cliprows) for whole drawFFT k-images, driven through
the reference's call pattern (SBProfile::drawK src/SBProfile.cpp:383-425, FillQuadrant 447-488,
SBExponentialImpl::fillKImage src/SBExponential.cpp:442-510).  The 2-d range test reads
std::abs(ky0+m+dky) (line 544), so rows are drawn that take a different branch under it than
under |ky0 + m dky|."""
import math

import numpy as np
import pytest

from oracle import profiles as pr
from oracle import ref

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")

KVA = 1e-5
KSQ_MAX = KVA ** (-1. / 1.5) - 1.          # SBExponentialImpl ctor (src/SBExponential.cpp:71-72)
K_MAX = math.sqrt(KSQ_MAX)


def _rows_2d(rng, n):
    """Random rows of a sheared k-image, plus rows built so that |ky0 + m dky| and
    |ky0 + m + dky| fall on opposite sides of kmax."""
    rows = []
    for _ in range(n):
        m = int(rng.integers(2, 400))
        rows.append((m, rng.uniform(-60, 60), rng.uniform(-2, 2), rng.uniform(-60, 60),
                     rng.uniform(-2, 2)))
    for _ in range(n):
        m = int(rng.integers(2, 30))
        dky = rng.uniform(0.5, 4.) * K_MAX / m
        ky0 = rng.uniform(-0.3, 0.3) * K_MAX
        dkx = rng.uniform(-0.9, 0.9) * K_MAX / m
        rows.append((m, rng.uniform(-0.2, 0.2) * K_MAX, dkx, ky0, dky))
    return rows


def test_oracle_range_2d_matches_reference():
    rng = np.random.default_rng(11)
    differs = 0
    for m, kx0, dkx, ky0, dky in _rows_2d(rng, 3000):
        got = pr._kvalue_range_2d(m, K_MAX, KSQ_MAX, kx0, dkx, ky0, dky)[:2]
        want = ref.kvalue_range_2d(m, K_MAX, KSQ_MAX, kx0, dkx, ky0, dky)
        assert tuple(got) == want, (m, kx0, dkx, ky0, dky, got, want)
        intended = (abs(ky0 + m * dky) > K_MAX) != (abs(ky0 + m + dky) > K_MAX)
        differs += int(intended)
    assert differs > 500      # the line-544 branch was exercised both ways


def test_oracle_range_1d_matches_reference():
    rng = np.random.default_rng(12)
    for _ in range(5000):
        m = int(rng.integers(1, 400))
        kx0 = rng.uniform(-60, 60) if rng.uniform() < 0.5 else 0.
        dkx = rng.uniform(0.01, 3.) * rng.choice([-1., 1.])
        ky = rng.uniform(-1.2, 1.2) * K_MAX
        got = pr._kvalue_range_1d(m, K_MAX, KSQ_MAX, kx0, dkx, ky)[:2]
        want = ref.kvalue_range_1d(m, K_MAX, KSQ_MAX, kx0, dkx, ky)
        assert tuple(got) == want, (m, kx0, dkx, ky, got, want)


def _reference_cliprows(r0, jac, Nk, dk):
    """The clip range of every row of the drawFFT k-image (x 0..Nk/2, y -Nk/2..Nk/2) as the
    reference's drawK + SBExponential fill computes them, calling its own range functions."""
    m = Nk // 2 + 1
    n = Nk + 1
    x0, y0 = 0., -(Nk // 2) * dk
    out = np.zeros((n, 2), dtype=np.int32)
    if jac is None or (jac[1] == 0. and jac[2] == 0.):
        mA, mD = (1., 1.) if jac is None else (jac[0], jac[3])
        # FillQuadrant with m1 = 0, n1 = Nk/2: a quadrant of rows ky = 0, dky, ... (accumulated),
        # row t of the image takes quadrant row |t|
        dkx, dky = dk * mA * r0, dk * mD * r0
        ky = 0.
        q = []
        for j in range(Nk // 2 + 1):
            q.append(ref.kvalue_range_1d(m, K_MAX, KSQ_MAX, (x0 * mA) * r0, dkx, ky))
            ky += dky
        for t in range(-(Nk // 2), Nk // 2 + 1):
            out[t + Nk // 2] = q[abs(t)]
        return out
    mA, mB, mC, mD = jac
    kx0 = (mA * x0 + mC * y0) * r0
    ky0 = (mB * x0 + mD * y0) * r0
    dkx, dkxy, dky, dkyx = mA * dk * r0, mC * dk * r0, mD * dk * r0, mB * dk * r0
    for j in range(n):
        out[j] = ref.kvalue_range_2d(m, K_MAX, KSQ_MAX, kx0, dkx, ky0, dkyx)
        kx0 += dkxy
        ky0 += dky
    return out


@pytest.mark.parametrize("jac", [None, (2.3, 0., 0., 1.7), (-1.9, 0., 0., 2.6),
                                 (2.1, 0.6, -0.4, 1.8), (1.2, -0.9, 0.7, 1.1)])
@pytest.mark.parametrize("Nk,r0", [(128, 12.), (256, 20.), (130, 11.), (768, 30.)])
def test_product_cliprows_match_reference(jac, Nk, r0):
    from galsim_amd import _lib
    dk = 2. * math.pi / Nk
    out = np.zeros(2 * (Nk + 1), dtype=np.int32)
    j = None if jac is None else np.array(jac, dtype=float)
    _lib.check(_lib.lib().gsmi_exponential_cliprows(r0, None, _lib.ptr(j), Nk, dk,
                                                    _lib.ptr(out)), "cliprows")
    want = _reference_cliprows(r0, jac, Nk, dk)
    got = out.reshape(-1, 2)
    # rows the reference leaves empty (i1 = i2 = m) may be stored as any empty range
    empty = want[:, 0] >= want[:, 1]
    assert np.array_equal(got[~empty], want[~empty]), (jac, Nk, r0)
    assert np.all(got[empty, 0] >= got[empty, 1])
    assert np.any(want[:, 1] - want[:, 0] < Nk // 2 + 1)     # some rows are clipped
