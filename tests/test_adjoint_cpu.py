"""CPU check of the adjoint's structure.

The adjoint kernels write dL/dtheta = sum over the stored k-image samples s of
Re[dF(s)/dtheta G(s)], where G(s) collects the seeds w_c Rhat(c, r) / N^2 of the wrapped cells
the sample feeds (conjugated where the forward adds conj F), Rhat = sum_p R(p) e^{+2 pi i k.p/N}
and R = dL/dI.  Here that sum, with dF/dtheta taken by central differences of the emulated
k-image, is compared with the central difference of the emulated image-level loss sum R I.  A
wrong Hermitian weight, a missing conjugate, the sign of the transform or a dropped alias fails
it.
"""
import numpy as np
import pytest

import galsim_amd as g
from galsim_amd import render
from galsim_amd.batch import _image_profile, _prepare_differentiable
from oracle import draw as od
from tests.program_emulator import eval_axis, eval_program

CASES = [
    ("exp_moffat", lambda m: m.Convolve([m.Exponential(half_light_radius=1.5, flux=1e4)
                                         .shear(g1=0.1, g2=-0.05), m.Moffat(beta=3, fwhm=0.7)]),
     (64, 64), (0.13, -0.21)),
    ("kolmogorov_folded", lambda m: m.Convolve([m.Gaussian(sigma=0.3).shear(g1=0.2, g2=0.1),
                                                m.Kolmogorov(fwhm=0.5)]), (32, 28), (0.2, 0.)),
    ("bulge_disk_shifted", lambda m: m.Convolve([
        (m.Exponential(half_light_radius=1.0, flux=600.) +
         m.Sersic(2.5, half_light_radius=0.5, flux=400.)).shear(g1=0.3, g2=-0.2).shift(0.1, 0.05),
        m.Kolmogorov(fwhm=0.75)]), (48, 40), (0.3, -0.4)),
]


class _Stamp:
    def __init__(self, make, shape, offset):
        nx, ny = shape
        p, b, _ = _image_profile(make(g), nx, ny, 0.2, offset, True, 'auto')
        self.N, self.Nk, self.dk = render.plan_sizes(p, b)
        ctx = render.Lowering('f64', self.Nk, self.dk)
        p._lower(ctx, None)
        nops, naxis = render.hoist_axis_factors(ctx, 0)
        naxis += _prepare_differentiable(ctx, 0, nops, naxis, 0)
        self.ctx, self.nops, self.b = ctx, nops, b
        self.prog = ctx.ops[:nops]
        self.axis = ctx.ops[nops:nops + naxis]
        hk = self.Nk // 2
        self.ii, self.jj = np.meshgrid(np.arange(0, hk + 1), np.arange(-hk, hk + 1))

    def kimage(self):
        ops = self.ctx.ops
        self.ctx.ops = self.prog
        try:
            F = eval_program(self.ctx, self.ii, self.jj, self.dk, self.Nk)
        finally:
            self.ctx.ops = ops
        return F * eval_axis(self.ctx, self.axis, self.ii, self.jj, self.dk)

    def image(self, F):
        N, b = self.N, self.b
        real = od.irfft_shifted(od.wrap_hermx(F, self.Nk, N), N)
        return real[b.ymin + N // 2: b.ymax + N // 2 + 1, b.xmin + N // 2: b.xmax + N // 2 + 1]

    def seeds(self, R):
        N, b = self.N, self.b
        Rfull = np.zeros((N, N))
        ys = np.arange(b.ymin, b.ymax + 1) % N
        xs = np.arange(b.xmin, b.xmax + 1) % N
        Rfull[np.ix_(ys, xs)] = R
        Rhat = np.fft.ifft2(Rfull) * N * N               # sum_p R(p) e^{+2 pi i k.p / N}
        w = np.full(N // 2 + 1, 2.)
        w[0] = w[-1] = 1.
        A = Rhat[:, :N // 2 + 1] * w / (N * N)            # A[r, c]
        ii, jj = self.ii, self.jj
        m = ii % N
        G = np.where(m <= N // 2, A[jj % N, np.minimum(m, N // 2)], 0.)
        mc = (-ii) % N
        conj = (ii > 0) & (mc <= N // 2)
        G = G + np.where(conj, np.conj(A[(-jj) % N, np.minimum(mc, N // 2)]), 0.)
        return G

    def get(self, f):
        kind, idx, name = f
        if kind == 'leaf':
            rec = self.ctx.leaves[idx]
            return rec['jac'][name] if isinstance(name, int) else float(rec[name])
        return self.axis[idx][{'a': 2, 'b': 3, 'c': 4}[name]]

    def set(self, f, v):
        kind, idx, name = f
        if kind == 'leaf':
            rec = self.ctx.leaves[idx]
            if isinstance(name, int):
                rec['jac'][name] = v
            else:
                rec[name] = v
        else:
            op = list(self.axis[idx])
            op[{'a': 2, 'b': 3, 'c': 4}[name]] = v
            self.axis[idx] = tuple(op)


def _fields(st):
    out = []
    for code, arg, *_ in st.prog:
        if code == g._lib.OP_LEAF and st.ctx.leaves[arg]['type'] != g._lib.LEAF_BOX:
            out += [('leaf', arg, 'flux'), ('leaf', arg, 'scale')]
            out += [('leaf', arg, k) for k in range(4)]
    ph = [k for k, op in enumerate(st.axis) if op[0] == g._lib.OP_PHASE][0]
    out += [('axis', ph, 'a'), ('axis', ph, 'b'), ('axis', ph, 'c')]
    return out


@pytest.mark.parametrize("name,make,shape,offset", CASES)
def test_adjoint_structure_matches_finite_differences(name, make, shape, offset):
    st = _Stamp(make, shape, offset)
    rng = np.random.default_rng(5)
    R = rng.normal(size=(shape[1], shape[0]))
    G = st.seeds(R)
    for f in _fields(st):
        v0 = st.get(f)
        h = 1e-4 * max(abs(v0), 1.)          # any step: both sides share Fp, Fm
        st.set(f, v0 + h)
        Fp = st.kimage()
        Lp = np.sum(R * st.image(Fp))
        st.set(f, v0 - h)
        Fm = st.kimage()
        Lm = np.sum(R * st.image(Fm))
        st.set(f, v0)
        fd = (Lp - Lm) / (2 * h)
        terms = np.real((Fp - Fm) / (2 * h) * G)
        adj = np.sum(terms)
        # the image is linear in the k-image, so the two agree to rounding for any dF
        assert abs(adj - fd) <= 1e-8 * max(np.sum(np.abs(terms)), 1e-300), (name, f, adj, fd)


def test_exp_moffat_field_map_matches_lowering():
    """The torch parameter -> field map of diff.ExpMoffat reproduces the records the host
    lowering writes for the same profile (checked at several parameter rows)."""
    import math
    import torch
    from galsim_amd import diff
    from galsim_amd.batch import BatchPlan  # noqa: F401  (lowering helpers live beside it)
    rng = np.random.default_rng(2)
    beta = 3.0
    for _ in range(5):
        th = np.array([rng.uniform(1e2, 1e5), rng.uniform(0.8, 2.5), rng.uniform(-0.3, 0.3),
                       rng.uniform(-0.3, 0.3), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5),
                       rng.uniform(0.6, 1.0)])
        prof = g.Convolve([g.Exponential(half_light_radius=th[1], flux=th[0])
                           .shear(g1=th[2], g2=th[3]), g.Moffat(beta, fwhm=th[6])])
        p, b, _ = _image_profile(prof, 64, 64, 0.2, th[4:6], True, 'auto')
        N, Nk, dk = render.plan_sizes(p, b)
        ctx = render.Lowering('f64', Nk, dk)
        p._lower(ctx, None)
        nops, naxis = render.hoist_axis_factors(ctx, 0)
        naxis += _prepare_differentiable(ctx, 0, nops, naxis, 0)
        e, m = ctx.leaves[0], ctx.leaves[1]
        ph = [op for op in ctx.ops[nops:nops + naxis] if op[0] == g._lib.OP_PHASE][0]
        P = diff.PER_LEAF
        base = torch.zeros(1, diff.NFIELD, dtype=torch.float64)
        base[0, P] = float(m['flux'])
        base[0, P + 2:P + 6] = torch.from_numpy(np.array(m['jac'], dtype=float))
        base[0, diff.F_FS] = ph[4]
        t = torch.tensor(th[None], dtype=torch.float64)
        cx = torch.tensor([ph[2] - th[4]])
        cy = torch.tensor([ph[3] - th[5]])
        f = diff.exp_moffat_fields(t, base, cx, cy, 0.2,
                                   1. / (2. * math.sqrt(2. ** (1. / beta) - 1.)))[0].numpy()
        np.testing.assert_allclose(f[0], e['flux'], rtol=1e-15)
        np.testing.assert_allclose(f[1], e['scale'], rtol=1e-15)
        np.testing.assert_allclose(f[2:6], np.array(e['jac']), rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(f[P + 1], m['scale'], rtol=1e-15)
        np.testing.assert_allclose(f[P + 2:P + 6], np.array(m['jac']), rtol=1e-15)
        np.testing.assert_allclose(f[diff.F_UX], ph[2], rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(f[diff.F_UY], ph[3], rtol=1e-14, atol=1e-15)


def test_moffat_beta_field_map_and_aux_match_lowering():
    """With beta as a parameter: the knorm and rD the torch map gives, and the auxiliary record
    values (k = 0 value, Temme constants), equal what the lowering writes for a Moffat put on
    the k^nu K_nu form at that beta."""
    import torch
    from galsim_amd import diff
    from galsim_amd.moffat import temme_constants
    for beta in (2.5, 2.73, 3.0, 3.5, 4.21):
        prof = g.Moffat(beta, fwhm=0.8)
        p, b, _ = _image_profile(prof, 40, 40, 0.2, (0., 0.), True, 'auto')
        N, Nk, dk = render.plan_sizes(p, b)
        ctx = render.Lowering('f64', Nk, dk)
        p._lower(ctx, None)
        nops, naxis = render.hoist_axis_factors(ctx, 0)
        _prepare_differentiable(ctx, 0, nops, naxis, 0, moffat_generic=True)
        rec = ctx.leaves[0]
        assert rec['kind'] == 0
        tb = torch.tensor([beta], dtype=torch.float64)
        aux = diff.moffat_aux(tb)[0].numpy()
        np.testing.assert_allclose(aux[0], rec['p'][1], rtol=1e-13)
        np.testing.assert_allclose(aux[1:], rec['p'][2:6], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(rec['p'][2:6], temme_constants(beta - 1.), rtol=1e-15)
        th = torch.tensor([[1., 1., 0., 0., 0., 0., 0.8]], dtype=torch.float64)
        base = torch.zeros(1, diff.NFIELD, dtype=torch.float64)
        f = diff.exp_moffat_fields(th, base, torch.zeros(1), torch.zeros(1), 0.2, None,
                                   beta=tb)[0].numpy()
        P = diff.PER_LEAF
        np.testing.assert_allclose(f[P], rec['flux'], rtol=1e-13)
        np.testing.assert_allclose(f[P + 1], rec['scale'], rtol=1e-14)
        assert f[P + diff.F_SHAPE] == beta
