"""Batched maximum-likelihood fits of stamp parameters on the GPU.

``BatchFit`` runs Adam on per-stamp parameters of a differentiable model (``diff.ExpMoffat``)
against noisy target stamps, minimising chi^2 = sum ((model - target) / sigma)^2.  One step --
parameters to fields (torch), the batched FP32/FP64 forward kernels, the chi^2 residual, the
adjoint kernels, the chain rule to the parameters, the Adam update -- is captured in a single
HIP graph and replayed.  The step contains no reduction over pixels: the chi^2 gradient is fed to
the adjoint directly as 2 (model - target) / sigma^2, so no library reduction (whose scratch
clearing could not be replayed safely) enters the graph.  Parameters are optimised in units of a
per-parameter scale (u = (theta - theta0) / scale), so one Adam learning rate suits flux, sizes,
shears and offsets alike.
"""
import numpy as np


class BatchFit:
    """Adam fit of ``model`` (a diff.ExpMoffat, or anything with ``render(theta)`` and ``dplan``)
    to ``target`` (flat tensor in the model's output layout) with per-pixel ``inv_sigma``."""

    def __init__(self, model, theta0, target, inv_sigma, scale, lr=0.05, betas=(0.9, 0.999)):
        import torch
        self.model = model
        self.theta0 = torch.as_tensor(theta0, dtype=torch.float64, device='cuda').clone()
        self.scale = torch.as_tensor(scale, dtype=torch.float64, device='cuda').reshape(1, -1)
        self.u = torch.zeros_like(self.theta0, requires_grad=True)
        self.target = target.detach().double()
        self.w2 = (2. * inv_sigma.double() ** 2).detach()
        self.opt = torch.optim.Adam([self.u], lr=lr, betas=betas, capturable=True)
        self._graph = None

    @property
    def theta(self):
        return (self.theta0 + self.u.detach() * self.scale)

    def _step(self):
        import torch
        self.opt.zero_grad(set_to_none=False)
        theta = self.theta0 + self.u * self.scale
        img = self.model.render(theta)
        dimg = (img.double() - self.target) * self.w2        # d chi^2 / d image
        torch.autograd.backward(img, grad_tensors=dimg.to(img.dtype))
        self.opt.step()

    def step(self):
        """One eager Adam step."""
        self._step()

    def capture(self, warmup=3):
        """Capture one step in a HIP graph (after ``warmup`` eager steps on a side stream, which
        also allocate Adam's state); ``replay`` then runs a step with one graph launch."""
        import torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._step()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._step()
        torch.cuda.synchronize()
        self._graph = g
        return g

    def replay(self, steps=1):
        for _ in range(steps):
            self._graph.replay()

    def chi2(self):
        """Per-stamp chi^2 at the current parameters (eager, outside any graph)."""
        import torch
        with torch.no_grad():
            r = (self.model.render(self.theta).double() - self.target) ** 2 * self.w2 / 2.
            return torch.stack([s.sum() for s in self.model.dplan.stamps(r)])


def exp_moffat_fit_problem(n, sizes=(32, 48, 64, 96, 128), seed=0, precision='f32',
                           snr=(20., 200.), beta=None, fit_beta=True):
    """The batched-fit workload: n stamps with sides drawn from ``sizes``, Exponential (x)
    Moffat with seeded true parameters (hlr ~ U(0.8, 2.5)", |g| ~ U(0, 0.5), sub-pixel offsets,
    flux 1e4, psf fwhm ~ U(0.6, 1.0)", beta ~ U(2.5, 4.5) unless given), noisy targets with a
    sky sigma giving S/N ~ U(snr) (S/N = flux / (sigma sqrt(4 pi r_eff^2)), r_eff the
    half-light radius in pixels of the PSF-convolved profile, approximately), and an initial
    guess perturbed from the truth.  With ``fit_beta`` the Moffat beta is the eighth parameter
    (initial guess off by up to 0.2); otherwise it is held at the truth.  Returns (model,
    theta_true, theta_init, target, inv_sigma, scale, betas, sides)."""
    import torch
    from . import diff
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.array(sizes), n)
    hlr = rng.uniform(0.8, 2.5, n)
    gm = rng.uniform(0., 0.5, n)
    ph = rng.uniform(0., np.pi, n)
    off = rng.uniform(-0.5, 0.5, (n, 2))
    fw = rng.uniform(0.6, 1.0, n)
    betas = rng.uniform(2.5, 4.5, n) if beta is None else np.full(n, float(beta))
    flux = np.full(n, 1e4)
    theta = np.stack([flux, hlr, gm * np.cos(2 * ph), gm * np.sin(2 * ph), off[:, 0],
                      off[:, 1], fw], axis=1)
    init = theta * np.stack([rng.uniform(0.9, 1.1, n), rng.uniform(0.9, 1.1, n),
                             np.ones(n), np.ones(n), np.ones(n), np.ones(n),
                             rng.uniform(0.95, 1.05, n)], axis=1)
    init[:, 2:4] += rng.normal(0., 0.03, (n, 2))
    init[:, 4:6] += rng.normal(0., 0.1, (n, 2))
    scale = np.array([1e3, 0.1, 0.05, 0.05, 0.2, 0.2, 0.05])
    if fit_beta:
        theta = np.concatenate([theta, betas[:, None]], axis=1)
        init = np.concatenate([init, (betas + rng.uniform(-0.2, 0.2, n))[:, None]], axis=1)
        scale = np.append(scale, 0.1)
    model = diff.ExpMoffat(init, sides, sides, scale=0.2, beta=betas, precision=precision,
                           fit_beta=fit_beta)
    t_true = torch.tensor(theta, dtype=torch.float64, device='cuda')
    with torch.no_grad():
        truth = diff.ExpMoffat(theta, sides, sides, scale=0.2, beta=betas, precision='f64',
                               fit_beta=fit_beta)
        timg = truth.render(t_true)
        # the truth's stamps in the fit model's layout (FFT sizes, hence order, may differ)
        img = torch.empty(model.dplan.plan.out.numel(), dtype=torch.float64, device='cuda')
        for a, b in zip(model.dplan.stamps(img), truth.dplan.stamps(timg)):
            a.copy_(b)
    snr_i = rng.uniform(snr[0], snr[1], n)
    r_pix = np.sqrt(hlr ** 2 + (fw / 2.) ** 2) / 0.2
    sky = flux / (snr_i * np.sqrt(4. * np.pi) * r_pix)
    inv_sigma = torch.empty_like(img)
    for k, s in enumerate(model.dplan.stamps(inv_sigma)):
        s.fill_(1. / sky[k])
    gen = torch.Generator(device='cpu').manual_seed(seed + 1)
    noise = torch.randn(img.shape, generator=gen, dtype=torch.float64).cuda()
    target = img + noise / inv_sigma
    return model, theta, init, target, inv_sigma, scale, betas, sides
