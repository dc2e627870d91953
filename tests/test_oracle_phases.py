"""Closed-form pins of the oracle's shift phases and offset composition (oracle/draw.py
apply_kimage_phases, _fwd_normal, Transformation._drawKImage), which restate ApplyKImagePhases
(src/SBTransform.cpp:740-883) and galsim/transform.py:192-230, 436-444, 592-600.

A transformed Gaussian has an exact Fourier transform.  If f'(x) = a f(J^-1 (x - x0)) then
F'(k) = a |det J| F(J^T k) e^{-i k.x0}, and a Gaussian of flux F and width s has
F(q) = F exp(-s^2 |q|^2 / 2).  Every case is compared with that closed form at 1e-12 of the
peak, and each pin is shown to catch a planted mistake in the oracle (a transposed _fwd_normal;
the composed Jacobian used for the phase instead of the outer one)."""
import math

import numpy as np
import pytest

from oracle import draw as od

TOL = 1e-12


def _shear(g1, g2):
    # galsim/shear.py:365-386 for a reduced shear (g1, g2)
    gsq = g1 * g1 + g2 * g2
    return np.array([[1. + g1, g2], [g2, 1. - g1]]) / math.sqrt(1. - gsq)


def _rot(t):
    return np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])


def _grid(nx, ny, dk):
    # the oracle's drawKImage grid: pixel i at kx = i dk, i = -nx//2 .. nx - 1 - nx//2
    kx = (np.arange(nx) - nx // 2) * dk
    ky = (np.arange(ny) - ny // 2) * dk
    return np.meshgrid(kx, ky)


def _gauss_k(u, separable):
    """SBGaussian's k-space value per unit flux at u = sigma^2 |q|^2 as its fills compute it
    (src/SBGaussian.cpp:200-290): exp(-u/2); the general-Jacobian fill takes
    1 - u/2 (1 - u/4) below ksq_min = (48 kvalue_accuracy)^(1/3) and 0 above
    ksq_max = -2 ln kvalue_accuracy (pinned by the leaf tests of the oracle)."""
    if separable:
        return np.exp(-0.5 * u)
    kva = 1e-5
    out = np.where(u < (kva * 48.) ** (1. / 3.), 1. - 0.5 * u * (1. - 0.25 * u), np.exp(-0.5 * u))
    return np.where(u > -2. * math.log(kva), 0., out)


def closed_form(terms, nx, ny, dk):
    """sum over terms (amp, flux, sigma, A, x0) of amp flux G(sigma^2 |A^T k|^2) exp(-i k.x0) on
    the drawKImage grid, G the Gaussian's radial function (diagonal A: the separable fill)."""
    KX, KY = _grid(nx, ny, dk)
    out = np.zeros(KX.shape, dtype=complex)
    for amp, flux, sig, A, x0 in terms:
        qx = A[0, 0] * KX + A[1, 0] * KY
        qy = A[0, 1] * KX + A[1, 1] * KY
        sep = A[0, 1] == 0. and A[1, 0] == 0.
        out += amp * flux * _gauss_k(sig * sig * (qx * qx + qy * qy), sep) * \
            np.exp(-1j * (KX * x0[0] + KY * x0[1]))
    return out


def _cases(m):
    """(name, profile built with module m, closed-form terms).  The profile constructors follow
    galsim's: shift(dx, dy), shear(g1=, g2=), rotate(theta), transform(dudx, dudy, dvdx, dvdy)."""
    S = _shear(0.2, -0.1)
    Sa = _shear(-0.15, 0.25)
    R = _rot(0.4)
    d = np.array([0.3, -0.2])
    d1 = np.array([-0.45, 0.35])
    Jd = np.diag([1.3, 0.7])
    T4 = np.array([[0.9, 0.3], [-0.2, 1.1]])

    def rotate(obj, t):
        if m is od:
            return obj.rotate(t)
        return obj.rotate(t * m.radians)

    def shear(obj, g1, g2):
        return obj.shear(g1=g1, g2=g2)

    G = lambda s, f: m.Gaussian(sigma=s, flux=f)
    return [
        # a shift inside a rotation inside a shear: composed by _fwd_normal (offset S R d);
        # drawn at the top level with no Jacobian (separable phase branch)
        ("shift_rotate_shear", shear(rotate(G(0.7, 2.).shift(*d), 0.4), 0.2, -0.1),
         [(1., 2., 0.7, S @ R, S @ R @ d)]),
        # a shifted, sheared component of a sum under a shear and a rotation: the phase takes
        # the outer Jacobian R S, the profile the composed one R S Sa (2-d phase branch)
        ("sum_inner_jac", rotate(shear(shear(G(0.6, 1.5), -0.15, 0.25).shift(*d1) + G(0.9, 0.8),
                                       0.2, -0.1), 0.4),
         [(1., 1.5, 0.6, R @ S @ Sa, R @ S @ d1), (1., 0.8, 0.9, R @ S, np.zeros(2))]),
        # a diagonal anisotropic Jacobian over a sum: separable phase branch with a Jacobian,
        # and the flux scaling |det J| multiplied without a phase
        ("sum_diag_aniso", (G(0.6, 1.5).shift(*d1) + G(0.9, 0.8)).transform(1.3, 0., 0., 0.7),
         [(0.91, 1.5, 0.6, Jd, Jd @ d1), (0.91, 0.8, 0.9, Jd, np.zeros(2))]),
        # a general Jacobian and a flux ratio composed onto a shift: fs = 2.5 |det T| with the
        # phase e^{-i k.T d}
        ("shift_general_flux", G(0.5, 1.2).shift(*d).transform(*T4.ravel()) * 2.5,
         [(2.5 * abs(np.linalg.det(T4)), 1.2, 0.5, T4, T4 @ d)]),
    ]


def _max_err(obj, terms, nx=41, ny=37, dk=0.21):
    got = obj.drawKImage(nx, ny, dk)
    ref = closed_form(terms, nx, ny, dk)
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


@pytest.mark.parametrize("idx", range(4), ids=[c[0] for c in _cases(od)])
def test_phases_match_closed_form(idx):
    name, obj, terms = _cases(od)[idx]
    err = _max_err(obj, terms)
    assert err < TOL, (name, err)


def test_planted_transposed_fwd_normal_is_caught(monkeypatch):
    good = od._fwd_normal

    def transposed(jac, x, y):
        return good(jac.T, x, y)
    monkeypatch.setattr(od, "_fwd_normal", transposed)
    errs = {name: _max_err(obj, terms) for name, obj, terms in _cases(od)}
    assert errs["shift_rotate_shear"] > 1e-3
    assert errs["shift_general_flux"] > 1e-3


def test_planted_composed_jacobian_for_the_phase_is_caught(monkeypatch):
    def wrong(self, image, jac):
        jac1 = self._jac if jac is None else jac if self._jac is None else jac.dot(self._jac)
        self.original._drawKImage(image, jac1)
        if self._has_offset:
            od.apply_kimage_phases(image, jac1, self._dx, self._dy, self._flux_scaling)
        elif abs(self._flux_scaling - 1.) > self.gsparams.kvalue_accuracy:
            image.array *= self._flux_scaling
    monkeypatch.setattr(od.Transformation, "_drawKImage", wrong)
    name, obj, terms = _cases(od)[1]
    assert _max_err(obj, terms) > 1e-3


def test_planted_phase_sign_is_caught(monkeypatch):
    good = od._fillphase_1d
    monkeypatch.setattr(od, "_fillphase_1d", lambda m, k, dk: np.conj(good(m, k, dk)))
    for name, obj, terms in _cases(od):
        assert _max_err(obj, terms) > 1e-3, name
