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
