"""GPU tests of the adjoint kernels (gsmi_adjoint_f32/f64) and the torch.autograd.Function
around them.

* Field level, FP64: dL/dfield from the adjoint against central differences of the FP64 forward
  kernels with the same records (L = sum W I for a fixed random W), for every live field of
  every leaf kind, clipped Exponentials, Moffat kinds, folded stamps and an in-program shift.
* Parameter level, FP64: autograd through the Exponential x Moffat model against central
  differences with step 1e-6 of the chi^2 loss, and against the oracle's FP64 drawImage
  differenced the same way.
* FP32 adjoint against the FP64 adjoint.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import galsim_amd as g  # noqa: E402
from galsim_amd import _lib  # noqa: E402
from galsim_amd import diff  # noqa: E402


def _stamp_losses(dplan, flat, W):
    return torch.stack([(s.double() * w).sum() for s, w in zip(dplan.stamps(flat), W)])


def _live_fields(dplan, i):
    out = []
    for l in range(int(dplan.nleaves[i])):
        if dplan.leaf_types[i, l] != _lib.LEAF_BOX:
            out += [diff.PER_LEAF * l + k for k in range(6)]
    return out + [diff.F_UX, diff.F_UY, diff.F_FS]


def _check_fields_fd(profiles, nx, ny, offsets, rtol=2e-6, seed=0):
    dplan = diff.DiffPlan(profiles, nx, ny, 0.2, offsets, precision='f64')
    f0 = dplan.fields0.clone()
    gen = torch.Generator(device='cpu').manual_seed(seed)
    W = [torch.randn(int(a), int(b), generator=gen, dtype=torch.float64).cuda()
         for a, b in dplan.plan.stamp_shape]
    flatW = torch.zeros(dplan.plan.out.numel(), dtype=torch.float64, device='cuda')
    for w, s in zip(W, dplan.stamps(flatW)):
        s.copy_(w)
    dplan.write_fields(f0)
    dplan.forward()
    grad = dplan.adjoint(flatW).cpu().numpy()
    grad[:, diff.F_FS] /= f0[:, diff.F_FS].cpu().numpy()
    nf = diff.NFIELD
    for f in range(nf):
        live = [i for i in range(dplan.n) if f in _live_fields(dplan, i)]
        if not live:
            continue
        v0 = f0[:, f].abs().clamp(min=1.0)
        fds = []
        for step in (1e-6, 1e-7):
            h = step * v0
            fp = f0.clone()
            fp[:, f] += h
            dplan.write_fields(fp)
            Lp = _stamp_losses(dplan, dplan.forward(), W)
            fm = f0.clone()
            fm[:, f] -= h
            dplan.write_fields(fm)
            Lm = _stamp_losses(dplan, dplan.forward(), W)
            fds.append(((Lp - Lm) / (2 * h)).cpu().numpy())
        for i in live:
            # tolerance relative to the stamp's largest field sensitivity (per unit of field)
            ref = max(abs(grad[i, k]) * max(abs(float(f0[i, k])), 1.)
                      for k in _live_fields(dplan, i))
            tol = rtol * ref / max(abs(float(f0[i, f])), 1.)
            # Sersic's table, its low-k series and its high-k fit meet with small jumps
            # (SersicInfo::kValue), so a difference step can straddle one: one of two steps
            # has to agree
            err = min(abs(grad[i, f] - fd[i]) for fd in fds)
            assert err <= tol, (i, f, grad[i, f], [fd[i] for fd in fds], tol)
    return dplan


def test_adjoint_fields_exp_moffat_f64():
    from galsim_amd import workloads
    p = workloads.exp_moffat_params(4, seed=3)
    profs = [workloads.exp_moffat_profile(g, p, i) for i in range(4)]
    _check_fields_fd(profs, 64, 64, p['offset'])


def test_adjoint_fields_bulge_disk_f64():
    from galsim_amd import workloads
    profs, nx, ny, offs, _ = workloads.build('bulge_disk', 5, seed=11)
    _check_fields_fd(profs, nx, ny, np.array([[0.1, -0.3]] * 5))


@pytest.mark.parametrize("beta", [1.5, 2.0, 2.5, 3.5, 4.0, 2.7, 1.45])
def test_adjoint_fields_moffat_kinds_f64(beta):
    profs = [g.Convolve([g.Gaussian(sigma=0.5, flux=50.).shear(g1=0.2, g2=-0.1),
                         g.Moffat(beta, fwhm=0.8)])]
    _check_fields_fd(profs, 40, 40, np.array([[0.2, 0.1]]))


def test_adjoint_fields_truncated_moffat_f64():
    # kV_trunc is a spline of its Hankel table on (k rD)^2: the adjoint takes its derivative;
    # the table (a function of trunc / rD) is held fixed, as the finite differences of the
    # fields hold it
    profs = [g.Convolve([g.Gaussian(sigma=0.5, flux=50.).shear(g1=0.2, g2=-0.1),
                         g.Moffat(2.5, fwhm=0.8, trunc=2.0)])]
    _check_fields_fd(profs, 40, 40, np.array([[0.2, 0.1]]))


def test_adjoint_fields_inprogram_shift_and_clip_f64():
    # a shift on one component stays in the program (its phase is not a field, but every leaf
    # gradient goes through it); a small Exponential is clipped by the reference's k-range rule
    profs = [g.Convolve([g.Exponential(scale_radius=0.05, flux=10.).shear(g1=0.4, g2=0.1)
                         + g.Gaussian(sigma=0.4, flux=5.).shift(0.3, -0.2),
                         g.Kolmogorov(fwhm=0.6)])]
    _check_fields_fd(profs, 32, 36, np.array([[0., 0.]]))


def _chi2_setup(n, seed, precision):
    """Noisy targets rendered at seeded true parameters (sigma = 1% of the peak), and a model
    whose grids are planned at perturbed parameters t1, where the gradient is taken."""
    from galsim_amd import workloads
    p = workloads.exp_moffat_params(n, seed=seed)
    theta = np.stack([p['flux'], p['hlr'], p['g1'], p['g2'], p['offset'][:, 0],
                      p['offset'][:, 1], p['psf_fwhm']], axis=1)
    truth = diff.ExpMoffat(theta, 64, 64, scale=0.2, beta=3.0, precision=precision)
    t = torch.tensor(theta, dtype=torch.float64, device='cuda')
    with torch.no_grad():
        img = truth.render(t).double()
    gen = torch.Generator(device='cpu').manual_seed(seed)
    sig = 0.01 * img.abs().max()
    target = img + sig * torch.randn(img.shape, generator=gen, dtype=torch.float64).cuda()
    t1 = t.clone()
    t1[:, 0] *= 1.05
    t1[:, 1] *= 0.97
    t1[:, 2] += 0.02
    t1[:, 4] += 0.1
    t1[:, 6] *= 1.03
    model = diff.ExpMoffat(t1.cpu().numpy(), 64, 64, scale=0.2, beta=3.0, precision=precision)
    return model, t1, target, sig, theta


def test_autograd_exp_moffat_matches_central_fd_f64():
    model, t1, target, sig, _ = _chi2_setup(3, 7, 'f64')

    def loss(t):
        return (((model.render(t).double() - target) / sig) ** 2).sum()

    t = t1.clone().requires_grad_(True)
    L = loss(t)
    L.backward()
    gr = t.grad.cpu().numpy()
    h = 1e-6
    with torch.no_grad():
        for k in range(7):
            tp = t1.clone()
            tp[:, k] += h
            tm = t1.clone()
            tm[:, k] -= h
            # per-stamp losses: stamps are independent, so perturb all at once
            ip = model.render(tp).double()
            im = model.render(tm).double()
            lp = (((ip - target) / sig) ** 2)
            lm = (((im - target) / sig) ** 2)
            st_p = torch.stack([s.sum() for s in model.dplan.stamps(lp)])
            st_m = torch.stack([s.sum() for s in model.dplan.stamps(lm)])
            fd = ((st_p - st_m) / (2 * h)).cpu().numpy()
            for i in range(len(fd)):
                scale = np.max(np.abs(gr[i]))
                assert abs(gr[i, k] - fd[i]) <= 1e-5 * max(abs(fd[i]), 1e-3 * scale), \
                    (i, k, gr[i, k], fd[i])


def test_autograd_exp_moffat_matches_oracle_fd():
    """Against FP64 central differences (step 1e-6) of the oracle's drawImage, the CPU
    restatement of GalSim's own drawImage(method='fft')."""
    from oracle import draw as od
    from galsim_amd import workloads
    model, t1, target, sig, _ = _chi2_setup(2, 9, 'f64')
    t = t1.clone().requires_grad_(True)
    L = (((model.render(t).double() - target) / sig) ** 2).sum()
    L.backward()
    gr = t.grad.cpu().numpy()
    tgt = [s.cpu().numpy() for s in model.dplan.stamps(target)]
    sg = float(sig)

    def oloss(i, th):
        p = dict(hlr=[th[1]], g1=[th[2]], g2=[th[3]], flux=[th[0]], beta=[3.0], psf_fwhm=[th[6]])
        prof = workloads.exp_moffat_profile(od, p, 0)
        im = prof.drawImage(64, 64, scale=0.2, offset=(th[4], th[5]))
        return float(np.sum(((im - tgt[i]) / sg) ** 2))

    th1 = t1.cpu().numpy()
    for i in range(2):
        scale = np.max(np.abs(gr[i]))
        for k in range(7):
            fds = []
            for step in (1e-6, 1e-7):
                # the oracle re-derives the Exponential's clipped k range at every call, so a
                # step can straddle a clip edge: one of two steps has to agree
                h = step * max(abs(th1[i, k]), 1.)
                tp = th1[i].copy()
                tp[k] += h
                tm = th1[i].copy()
                tm[k] -= h
                fds.append((oloss(i, tp) - oloss(i, tm)) / (2 * h))
            err = min(abs(gr[i, k] - fd) for fd in fds)
            assert err <= 2e-5 * max(abs(fds[0]), 1e-3 * scale), (i, k, gr[i, k], fds)


def test_fp32_adjoint_matches_fp64():
    from galsim_amd import workloads
    for kind, n in (('exp_moffat', 16), ('bulge_disk', 16)):
        profs, nx, ny, offs, _ = workloads.build(kind, n, seed=21)
        d64 = diff.DiffPlan(profs, nx, ny, 0.2, offs, precision='f64')
        d32 = diff.DiffPlan(profs, nx, ny, 0.2, offs, precision='f32')
        out64 = d64.forward().clone()
        gen = torch.Generator(device='cpu').manual_seed(1)
        W = torch.randn(out64.numel(), generator=gen, dtype=torch.float64).cuda()
        # a residual that looks like a chi^2 gradient: model minus a perturbed copy
        R = (out64 - out64 * (1 + 0.01 * W)) / out64.abs().max()
        g64 = d64.adjoint(R).cpu().numpy()
        g32 = d32.adjoint(R.float()).cpu().numpy()
        f0 = d64.fields0.cpu().numpy()
        worst = 0.
        for i in range(n):
            live = _live_fields(d64, i)
            s = np.abs(g64[i, live]) * np.maximum(np.abs(f0[i, live]), 1.)
            err = np.abs(g32[i, live] - g64[i, live]) * np.maximum(np.abs(f0[i, live]), 1.)
            worst = max(worst, np.max(err) / np.max(s))
        print("FP32 vs FP64 adjoint, %s: worst field error / largest sensitivity = %.2e"
              % (kind, worst))
        assert worst <= 3e-4, (kind, worst)


def test_adjoint_rejects_bad_residual_size():
    profs = [g.Gaussian(sigma=0.5)]
    d = diff.DiffPlan(profs, 16, 16, 0.2, None, precision='f32')
    with pytest.raises(ValueError):
        d.adjoint(torch.zeros(10, device='cuda'))


def test_autograd_beta_gradient_matches_central_fd_f64():
    """The Moffat beta as an eighth parameter (k^nu K_nu form, shape gradients in the adjoint):
    autograd against central differences of the chi^2 loss, every parameter."""
    from galsim_amd import workloads
    p = workloads.exp_moffat_params(3, seed=13)
    beta = np.array([2.7, 3.0, 4.2])
    theta = np.stack([p['flux'], p['hlr'], p['g1'], p['g2'], p['offset'][:, 0],
                      p['offset'][:, 1], p['psf_fwhm'], beta], axis=1)
    model = diff.ExpMoffat(theta, 48, 48, scale=0.2, precision='f64', fit_beta=True)
    t0 = torch.tensor(theta, dtype=torch.float64, device='cuda')
    with torch.no_grad():
        img = model.render(t0).double()
    gen = torch.Generator(device='cpu').manual_seed(2)
    sig = 0.01 * img.abs().max()
    target = img + sig * torch.randn(img.shape, generator=gen, dtype=torch.float64).cuda()
    t1 = t0.clone()
    t1[:, 1] *= 0.97
    t1[:, 6] *= 1.04
    t1[:, 7] += 0.15

    def losses(t):
        r = ((model.render(t).double() - target) / sig) ** 2
        return torch.stack([s.sum() for s in model.dplan.stamps(r)])

    t = t1.clone().requires_grad_(True)
    losses(t).sum().backward()
    gr = t.grad.cpu().numpy()
    with torch.no_grad():
        for k in range(8):
            h = 1e-6 * max(1., float(t1[:, k].abs().max()))
            tp = t1.clone()
            tp[:, k] += h
            tm = t1.clone()
            tm[:, k] -= h
            fd = ((losses(tp) - losses(tm)) / (2 * h)).cpu().numpy()
            for i in range(3):
                scale = np.max(np.abs(gr[i]))
                assert abs(gr[i, k] - fd[i]) <= 2e-5 * max(abs(fd[i]), 1e-3 * scale), \
                    (i, k, gr[i, k], fd[i])


def test_fp32_beta_gradient_matches_fp64():
    """The FP32 shape gradient (Moffat k^nu K_nu differentiated in nu by quadrature) against the
    FP64 one (differences of FP64 Bessel functions, itself checked against central differences
    of the loss above), with every other field, for beta from 1.45 to 7 and stamps of 32 to 128
    pixels, on the same residual."""
    from galsim_amd import workloads
    p = workloads.exp_moffat_params(7, seed=17)
    p['beta'] = np.array([1.45, 2.0, 2.7, 3.0, 4.2, 5.5, 7.0])
    profs = [workloads.exp_moffat_profile(g, p, i) for i in range(7)]
    sides = np.array([32, 48, 64, 96, 128, 64, 48])
    kw = dict(moffat_generic=True, shape_gradients=True)
    d64 = diff.DiffPlan(profs, sides, sides, 0.2, p['offset'], precision='f64', **kw)
    d32 = diff.DiffPlan(profs, sides, sides, 0.2, p['offset'], precision='f32', **kw)
    out64 = d64.forward().clone()
    gen = torch.Generator(device='cpu').manual_seed(3)
    W = torch.randn(out64.numel(), generator=gen, dtype=torch.float64).cuda()
    R = (out64 - out64 * (1 + 0.01 * W)) / out64.abs().max()
    g64 = d64.adjoint(R).cpu().numpy()
    g32 = d32.adjoint(R.float()).cpu().numpy()
    f0 = d64.fields0.cpu().numpy()
    shape = diff.PER_LEAF * 1 + diff.F_SHAPE           # the Moffat slot's beta field
    worst = worst_shape = 0.
    for i in range(7):
        live = _live_fields(d64, i) + [shape]
        s = np.abs(g64[i, live]) * np.maximum(np.abs(f0[i, live]), 1.)
        err = np.abs(g32[i, live] - g64[i, live]) * np.maximum(np.abs(f0[i, live]), 1.)
        worst = max(worst, np.max(err) / np.max(s))
        worst_shape = max(worst_shape, err[-1] / np.max(s))
        assert abs(g64[i, shape]) > 0., i
    print("FP32 vs FP64 adjoint with beta: worst field error / largest sensitivity = %.2e "
          "(beta field %.2e)" % (worst, worst_shape))
    assert worst <= 3e-4, worst


@pytest.mark.parametrize("fit_beta", [False, True])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_parameter_map_matches_torch_expressions(fit_beta, precision):
    """ExpMoffat's device parameter map and chain rule (gsmi_exp_moffat_fields / _grad) against
    the same model with the map written as torch expressions under torch's autograd: equal
    images and equal gradients of a chi^2 loss, every parameter (beta included)."""
    from galsim_amd import workloads
    p = workloads.exp_moffat_params(6, seed=29)
    beta = np.array([2.5, 3.0, 3.3, 4.1, 2.8, 3.7])
    cols = [p['flux'], p['hlr'], p['g1'], p['g2'], p['offset'][:, 0], p['offset'][:, 1],
            p['psf_fwhm']]
    if fit_beta:
        cols.append(beta)
    theta = np.stack(cols, axis=1)
    sides = np.array([32, 48, 64, 48, 96, 64])
    kw = dict(scale=0.2, beta=beta, precision=precision, fit_beta=fit_beta)
    fused = diff.ExpMoffat(theta, sides, sides, fused=True, **kw)
    plain = diff.ExpMoffat(theta, sides, sides, fused=False, **kw)
    t = torch.tensor(theta, dtype=torch.float64, device='cuda')
    t1 = t.clone()
    t1[:, 1] *= 1.05
    t1[:, 2] += 0.02
    t1[:, 6] *= 0.97
    if fit_beta:
        t1[:, 7] += 0.1
    with torch.no_grad():
        target = plain.render(t).double()
    grads, imgs = [], []
    for m in (fused, plain):
        tt = t1.clone().requires_grad_(True)
        img = m.render(tt)
        ((img.double() - target) ** 2).sum().backward()
        grads.append(tt.grad.cpu().numpy())
        imgs.append(img.detach().double().cpu().numpy())
    tol = 1e-12 if precision == 'f64' else 1e-6
    assert np.max(np.abs(imgs[0] - imgs[1])) <= tol * np.max(np.abs(imgs[1]))
    # per stamp, each gradient times its parameter's size (the loss change for a relative
    # step), against the stamp's largest such sensitivity
    w = np.maximum(np.abs(theta), 0.1)
    sens = np.max(np.abs(grads[1]) * w, axis=1, keepdims=True)
    err = np.abs(grads[0] - grads[1]) * w / sens
    print("fused vs torch parameter map: worst gradient difference %.2e of the largest "
          "sensitivity" % err.max())
    assert err.max() <= (1e-9 if precision == 'f64' else 1e-4), err
