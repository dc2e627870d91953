"""GPU tests of the batched Adam fit: the step replayed from a HIP graph equals the same steps
run eagerly, and the fit moves every parameter toward the truth and chi^2 toward the number of
pixels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from galsim_amd import fit  # noqa: E402


def _problem(n, seed, precision='f32', snr=(20., 200.)):
    return fit.exp_moffat_fit_problem(n, sizes=(32, 48, 64), seed=seed, precision=precision,
                                      snr=snr)


def test_graph_steps_equal_eager_steps():
    model, theta, init, target, isig, scale, betas, sides = _problem(6, 3)
    a = fit.BatchFit(model, init, target, isig, scale)
    for _ in range(5):
        a.step()
    ta = a.theta.clone()
    model2, *_ = _problem(6, 3)
    b = fit.BatchFit(model2, init, target, isig, scale)
    b.capture(warmup=3)
    b.replay(2)
    torch.cuda.synchronize()
    assert torch.allclose(b.theta, ta, rtol=1e-12, atol=1e-12), (b.theta - ta).abs().max()


def test_fit_converges_toward_truth():
    n = 12
    # high S/N, so that the starting guess is many noise sigmas from the optimum
    model, theta, init, target, isig, scale, betas, sides = _problem(n, 5, snr=(400., 800.))
    f = fit.BatchFit(model, init, target, isig, scale)
    chi0 = f.chi2().cpu().numpy()
    f.capture()
    f.replay(100)
    torch.cuda.synchronize()
    chi1 = f.chi2().cpu().numpy()
    npix = sides.astype(float) ** 2
    est = f.theta.cpu().numpy()
    print("chi2/npix before", np.round(chi0 / npix, 2), "after", np.round(chi1 / npix, 3))
    assert np.all(chi1 < chi0)
    assert np.all(chi1 / npix < 1.3)
    # each parameter closer to the truth than the initial guess, in units of its scale
    e0 = np.abs(init - theta) / scale
    e1 = np.abs(est - theta) / scale
    assert np.median(e1) < 0.5 * np.median(e0), (np.median(e0), np.median(e1))


def test_config4_workload_fp32_fit_matches_fp64_fit_and_oracle():
    # The batched-fit workload at its stated size: 8192 Exponential x Moffat stamps, sides in
    # {32, 48, 64, 96, 128}, beta ~ U(2.5, 4.5), S/N ~ U(20, 200), 50 Adam steps on all 8
    # parameters (flux, hlr, g1, g2, dx, dy, psf_fwhm, beta) replayed from a HIP graph.
    from galsim_amd import diff
    from oracle import draw as od
    n, steps, seed = 8192, 50, 11
    sizes = (32, 48, 64, 96, 128)
    fits = {}
    for prec in ('f32', 'f64'):
        print("config4 fit", prec, n, "stamps", flush=True)
        model, theta, init, target, isig, scale, betas, sides = fit.exp_moffat_fit_problem(
            n, sizes=sizes, seed=seed, precision=prec)
        assert theta.shape == (n, 8) and set(np.unique(sides)) == set(sizes)
        assert betas.min() >= 2.5 and betas.max() <= 4.5
        f = fit.BatchFit(model, init, target, isig, scale)
        f.capture()
        f.replay(steps)
        torch.cuda.synchronize()
        fits[prec] = (f.theta.cpu().numpy(), f.chi2().cpu().numpy() / sides.astype(float) ** 2)
    d = np.abs(fits['f32'][0] - fits['f64'][0]) / scale
    print("FP32 vs FP64 fit, max |diff| / scale per parameter:", np.round(d.max(axis=0), 7))
    # the FP32 fit lands on the FP64 fit's parameters to 1e-3 of each parameter's scale
    assert d.max() < 1e-3
    assert 0.8 < np.median(fits['f32'][1]) < 1.2 and 0.8 < np.median(fits['f64'][1]) < 1.2
    # renders at the fitted parameters against the oracle's drawImage: 6 stamps of each side
    est = fits['f64'][0]
    pick = np.concatenate([np.nonzero(sides == s)[0][:6] for s in sizes])
    t = est[pick]
    for prec, tol in (('f64', 1e-10), ('f32', 1e-5)):
        m = diff.ExpMoffat(t, sides[pick], sides[pick], scale=0.2, beta=t[:, 7], precision=prec,
                           fit_beta=True)
        with torch.no_grad():
            img = m.render(torch.tensor(t, dtype=torch.float64, device='cuda'))
        worst = 0.
        for k, stamp in enumerate(m.dplan.stamps(img)):
            p = t[k]
            gal = od.Exponential(half_light_radius=p[1], flux=p[0]).shear(g1=p[2], g2=p[3])
            ref = od.Convolve([gal, od.Moffat(beta=p[7], fwhm=p[6])]).drawImage(
                int(sides[pick][k]), int(sides[pick][k]), scale=0.2, offset=(p[4], p[5]))
            a = stamp.double().cpu().numpy()
            worst = max(worst, float(np.max(np.abs(a - ref)) / np.max(np.abs(ref))))
        print(prec, "renders at the fitted parameters, max peak-relative error vs oracle", worst)
        assert worst < tol
