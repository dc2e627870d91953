"""Differentiable batched rendering: gradients of a pixel loss through drawImage(method='fft').

``DiffPlan`` is a BatchPlan whose stamp fields -- per program leaf its flux, scale, composite
Jacobian and shape (the Moffat beta), and per stamp the phase (ux, uy) and amplitude fs applying
to the whole k-image -- can be rewritten on the device and differentiated.  ``render_fields`` is
a ``torch.autograd.Function`` over those fields: the forward runs the batched FFT drawing
kernels, the backward runs the hand-written adjoint kernels (gsmi_adjoint_f32/f64), which
recompute the k-space products rather than keeping any complex grid from the forward.

Model parameters map to fields through plain torch code (so torch's autograd does the outer
chain rule); ``ExpMoffat`` does this for Convolve(Exponential(hlr, flux).shear(g1, g2),
Moffat(beta, fwhm)) drawn with offset (dx, dy) -- the 7 parameters (flux, hlr, g1, g2, dx, dy,
psf_fwhm), and with ``fit_beta`` the Moffat beta as an eighth.

The FFT grids (N, Nk per stamp), the Sersic tables, and the Exponential clip rows of the FP64
path are those of the profiles the plan was built from: gradients are those of the rendering on
that fixed grid (the reference's size choices are piecewise constant in the parameters).
"""
import math

import numpy as np

from . import _lib
from .batch import BatchPlan

NFIELD = _lib.GRAD_FIELDS
MAXLEAF = _lib.GRAD_MAXLEAF
PER_LEAF = _lib.GRAD_PER_LEAF
MAXOPS = 12
F_UX, F_UY, F_FS = PER_LEAF * MAXLEAF, PER_LEAF * MAXLEAF + 1, PER_LEAF * MAXLEAF + 2
F_SHAPE = 6                                   # offset of the shape field within a leaf slot
_LEAF_FIELDS = (('flux', 0), ('scale', 0), ('jac', 0), ('jac', 1), ('jac', 2), ('jac', 3),
                ('p', 0))


class DiffPlan:
    """A device-resident, differentiable batch of stamps (see the module docstring).

    ``moffat_generic`` puts every Moffat leaf on the k^nu K_nu form so its beta can move;
    ``shape_gradients`` makes the adjoint compute the Moffat beta derivatives."""

    def __init__(self, profiles, nx, ny, scale, offsets=None, precision='f32',
                 use_true_center=True, moffat_generic=False, shape_gradients=False):
        import torch
        # a differentiable batch maps fields to every stamp: a stamp that cannot be planned
        # stops it (no stamps are left out)
        self.plan = BatchPlan(profiles, nx, ny, scale, offsets, precision, 'auto',
                              use_true_center, differentiable=True,
                              moffat_generic=moffat_generic, on_error='raise')
        p = self.plan
        n = p.n
        self.n = n
        self.precision = precision
        self.shape_gradients = bool(shape_gradients)
        recs = p.recs                       # sorted by N
        ops, leaves = p.host_ops, p.host_leaves
        o_ops, o_leaves = p.offsets['ops'], p.offsets['leaves']
        lf, of = _lib.LEAF_DTYPE.fields, _lib.OP_DTYPE.fields
        dst = np.full((n, NFIELD), -1, dtype=np.int64)
        f0 = np.zeros((n, NFIELD))
        self.leaf_types = np.zeros((n, MAXLEAF), dtype=np.int32)
        self._leaf_index = np.full((n, MAXLEAF), -1, dtype=np.int64)
        nl = np.zeros(n, dtype=np.int32)
        cplx = np.zeros(n, dtype=bool)
        inv = np.empty(n, dtype=np.int64)
        inv[p.order] = np.arange(n)
        for s in range(n):                  # s: sorted position
            r = recs[s]
            i = p.order[s]                  # input index
            b, cnt = int(r['op_begin']), int(r['op_count'])
            prog = ops[b:b + cnt]
            args = prog['arg'][prog['code'] == _lib.OP_LEAF]
            nleaf = int(args.max()) + 1
            if nleaf > MAXLEAF or cnt > MAXOPS:
                raise _lib.GalSimAMDError(
                    "differentiable rendering supports at most %d leaves and %d ops per profile "
                    "(got %d and %d)" % (MAXLEAF, MAXOPS, nleaf, cnt))
            lb = int(r['leaf_begin'])
            if np.any(leaves['type'][lb:lb + nleaf] == _lib.LEAF_INTERP_IMAGE):
                raise NotImplementedError("differentiable rendering of an InterpolatedImage is "
                                          "not part of galsim_amd")
            nl[i] = nleaf
            cplx[i] = bool(np.any(prog['code'] == _lib.OP_PHASE))
            for l in range(nleaf):
                li = int(r['leaf_begin']) + l
                rec = leaves[li]
                self.leaf_types[i, l] = rec['type']
                self._leaf_index[i, l] = li
                for f, (name, k) in enumerate(_LEAF_FIELDS):
                    byte = o_leaves + li * _lib.LEAF_DTYPE.itemsize + lf[name][1] + 8 * k
                    dst[i, PER_LEAF * l + f] = byte // 8
                    f0[i, PER_LEAF * l + f] = rec[name][k] if name in ('jac', 'p') else rec[name]
            axis = range(b + cnt, b + cnt + int(r['axis_count']))
            ph = [k for k in axis if ops[k]['code'] == _lib.OP_PHASE][0]
            for f, name in ((F_UX, 'a'), (F_UY, 'b'), (F_FS, 'c')):
                dst[i, f] = (o_ops + ph * _lib.OP_DTYPE.itemsize + of[name][1]) // 8
                f0[i, f] = ops[ph][name]
        self.nleaves = nl
        self._inv = torch.from_numpy(inv).cuda()          # input index -> sorted position
        mask = dst.reshape(-1) >= 0
        self._dst_table = torch.from_numpy(dst).cuda()          # (n, NFIELD), -1: none
        self._src = torch.from_numpy(np.nonzero(mask)[0]).cuda()
        self._dst = torch.from_numpy(dst.reshape(-1)[mask]).cuda()
        self.fields0 = torch.from_numpy(f0).cuda()
        self._dev64 = p.dev.view(torch.float64)
        self._aux_dst = None
        lib = _lib.lib()
        need = 1
        self._bins = []
        for N, first, count, max_ny in p.groups:
            nd = np.zeros(1, dtype=np.int64)
            _lib.check(lib.gsmi_adjoint_partials(int(N), int(count), _lib.ptr(nd)),
                       "adjoint partials size")
            need = max(need, int(nd[0]))
            members = p.order[first:first + count]
            flags = (_lib.ADJ_COMPLEX if cplx[members].any() else 0) | \
                (_lib.ADJ_SHAPE if self.shape_gradients else 0) | \
                (p.program_shape << 8)
            self._bins.append((N, first, count, max_ny, int(nl[members].max()), flags))
        self._partials = torch.empty(need, dtype=torch.float64, device='cuda')
        self._grad = torch.empty((n, NFIELD), dtype=torch.float64, device='cuda')
        # the forward as one fold and one row launch over work lists (specialised to the
        # batch's program shape when it has one)
        p.use_schedule()

    # -- fields <-> device records -------------------------------------------------------------
    def use_aux(self, slot, first, count):
        """Declare record values written alongside the fields but not differentiated: doubles
        p[first .. first + count) of leaf ``slot`` of every stamp (e.g. the Moffat constants that
        follow its beta)."""
        import torch
        lf = _lib.LEAF_DTYPE.fields
        li = self._leaf_index[:, slot]
        if np.any(li < 0):
            raise ValueError("leaf slot %d is not used by every stamp" % slot)
        base = self.plan.offsets['leaves'] + li * _lib.LEAF_DTYPE.itemsize + lf['p'][1]
        idx = base[:, None] // 8 + first + np.arange(count)[None, :]
        self._aux_dst = torch.from_numpy(np.ascontiguousarray(idx.reshape(-1))).cuda()

    def write_fields(self, fields, aux=None):
        """Copy a (n, NFIELD) float64 device tensor of fields (input order), and the declared
        auxiliary values ((n, count), see use_aux), into the records."""
        import torch
        if fields.shape != (self.n, NFIELD) or fields.dtype != torch.float64:
            raise ValueError("fields must be a (%d, %d) float64 tensor" % (self.n, NFIELD))
        vals = fields.detach().reshape(-1).index_select(0, self._src)
        self._dev64.index_copy_(0, self._dst, vals)
        if aux is not None:
            self._dev64.index_copy_(0, self._aux_dst, aux.detach().reshape(-1).to(torch.float64))

    def forward(self):
        """Run the drawing kernels on the current fields; returns the flat output tensor (the
        plan's own buffer: clone it before the next run if it must be kept)."""
        return self.plan.run()

    def adjoint(self, resid):
        """dL/dfield for every stamp (input order, (n, NFIELD) float64) given resid = dL/dI as a
        flat tensor laid out like the plan's output.  Field F_FS holds d/dlog(fs)."""
        return self.adjoint_sorted(resid).index_select(0, self._inv)

    def adjoint_sorted(self, resid):
        """adjoint() in the plan's sorted stamp order (row self._inv[i] is stamp i); the tensor
        is the plan's own buffer, overwritten by the next call."""
        import torch
        p = self.plan
        fdt = torch.float32 if self.precision == 'f32' else torch.float64
        resid = resid.to(fdt).contiguous()
        if resid.numel() != p.out.numel():
            raise ValueError("residual has %d values, the batch has %d pixels"
                             % (resid.numel(), p.out.numel()))
        lib = _lib.lib()
        base = p.dev.data_ptr()
        o = p.offsets
        s = torch.cuda.current_stream().cuda_stream
        g = self._grad
        for N, first, count, max_ny, ml, flags in self._bins:
            st = base + o['stamp'] + first * _lib.STAMP_DTYPE.itemsize
            gptr = g.data_ptr() + first * NFIELD * 8
            if self.precision == 'f32':
                rc = lib.gsmi_adjoint_f32(st, count, N, max_ny, ml, flags, base + o['ops'],
                                          base + o['leaves'], base + o['tables'],
                                          base + o['tables32'], resid.data_ptr(),
                                          p.scratch.data_ptr(), self._partials.data_ptr(),
                                          gptr, s)
            else:
                rc = lib.gsmi_adjoint_f64(st, count, N, max_ny, ml, flags, base + o['ops'],
                                          base + o['leaves'], base + o['tables'],
                                          base + o['clip'], resid.data_ptr(),
                                          p.scratch.data_ptr(), self._partials.data_ptr(),
                                          gptr, s)
            _lib.check(rc, "adjoint (N=%d)" % N)
        return g

    def stamps(self, flat):
        """Split a flat output-layout tensor into per-stamp (ny, nx) views (input order)."""
        out = []
        for i in range(self.n):
            off = int(self.plan.stamp_offset[i])
            ny, nx = (int(v) for v in self.plan.stamp_shape[i])
            out.append(flat[off:off + ny * nx].view(ny, nx))
        return out


def _function():
    import torch

    class RenderFields(torch.autograd.Function):
        @staticmethod
        def forward(ctx, fields, dplan, aux):
            dplan.write_fields(fields, aux)
            out = dplan.forward().clone()
            ctx.dplan = dplan
            ctx.aux = aux
            ctx.save_for_backward(fields)
            return out

        @staticmethod
        def backward(ctx, gout):
            fields, = ctx.saved_tensors
            dplan = ctx.dplan
            dplan.write_fields(fields, ctx.aux)   # another forward may have rewritten them
            g = dplan.adjoint(gout).clone()
            g[:, F_FS] = g[:, F_FS] / fields[:, F_FS]
            return g, None, None

    return RenderFields


_RenderFields = None


def render_fields(fields, dplan, aux=None):
    """Differentiable render: flat output tensor of ``dplan``'s stamps for the given fields
    (and auxiliary record values, not differentiated)."""
    global _RenderFields
    if _RenderFields is None:
        _RenderFields = _function()
    return _RenderFields.apply(fields, dplan, aux)


class ExpMoffat:
    """Convolve(Exponential(half_light_radius=hlr, flux=flux).shear(g1=g1, g2=g2),
    Moffat(beta, fwhm=psf_fwhm)) drawn by drawImage(nx, ny, scale=scale, offset=(dx, dy)) for a
    batch of parameter rows theta = (flux, hlr, g1, g2, dx, dy, psf_fwhm), differentiable in all
    seven; with ``fit_beta`` theta has an eighth column, beta, differentiable too.  ``beta``,
    ``nx`` and ``ny`` may be per stamp.  The grids are fixed by ``theta0``."""

    PARAMS = ('flux', 'hlr', 'g1', 'g2', 'dx', 'dy', 'psf_fwhm')

    def __init__(self, theta0, nx, ny, scale=0.2, beta=3.0, precision='f32', fit_beta=False,
                 fused=True):
        import torch
        from .convolve import Convolve
        from .exponential import Exponential
        from .moffat import Moffat
        self.fit_beta = bool(fit_beta)
        ncol = 8 if self.fit_beta else 7
        theta0 = np.asarray(theta0, dtype=float).reshape(-1, ncol)
        n = len(theta0)
        self.scale = float(scale)
        self.beta = (theta0[:, 7].copy() if self.fit_beta else
                     np.broadcast_to(np.asarray(beta, dtype=float), (n,)).copy())
        profs = []
        for t, b in zip(theta0, self.beta):
            gal = Exponential(half_light_radius=t[1], flux=t[0]).shear(g1=t[2], g2=t[3])
            profs.append(Convolve([gal, Moffat(float(b), fwhm=t[6])]))
        self.dplan = DiffPlan(profs, nx, ny, scale, theta0[:, 4:6], precision,
                              moffat_generic=self.fit_beta, shape_gradients=self.fit_beta)
        lt = self.dplan.leaf_types
        if not (np.all(lt[:, 0] == _lib.LEAF_EXPONENTIAL) and np.all(lt[:, 1] == _lib.LEAF_MOFFAT)):
            raise _lib.GalSimAMDError("unexpected lowering of the Exponential x Moffat model")
        if self.fit_beta:
            self.dplan.use_aux(1, 1, 5)          # Moffat p[1..5]: k = 0 value, Temme constants
        f0 = self.dplan.fields0
        t0 = torch.from_numpy(theta0).cuda()
        self._base = f0
        self._cx = f0[:, F_UX] - t0[:, 4]
        self._cy = f0[:, F_UY] - t0[:, 5]
        self._fwhm_to_rd = torch.from_numpy(
            1. / (2. * np.sqrt(2. ** (1. / self.beta) - 1.))).cuda()
        self._cxcy = torch.stack([self._cx, self._cy], dim=1).contiguous()
        self.theta0 = t0
        self.fused = bool(fused)

    def fields(self, theta):
        """(n, NFIELD) fields of parameter rows theta ((n, 7 or 8) float64 device tensor)."""
        if self.fit_beta:
            return exp_moffat_fields(theta[:, :7], self._base, self._cx, self._cy, self.scale,
                                     None, beta=theta[:, 7])
        return exp_moffat_fields(theta, self._base, self._cx, self._cy, self.scale,
                                 self._fwhm_to_rd)

    def render(self, theta):
        """Flat FP32/FP64 tensor of all stamps for theta; differentiable.  With ``fused`` (the
        default) the parameter map and its chain rule are one device kernel each
        (gsmi_exp_moffat_fields / _grad); otherwise they are the torch expressions of
        ``fields`` under torch's autograd."""
        if self.fused:
            global _ExpMoffatRender
            if _ExpMoffatRender is None:
                _ExpMoffatRender = _exp_moffat_function()
            return _ExpMoffatRender.apply(theta, self)
        aux = moffat_aux(theta[:, 7].detach()) if self.fit_beta else None
        return render_fields(self.fields(theta), self.dplan, aux)

    def _check_theta(self, theta):
        import torch
        ncol = 8 if self.fit_beta else 7
        if theta.shape != (self.dplan.n, ncol) or theta.dtype != torch.float64 or \
                not theta.is_cuda:
            raise ValueError("theta must be a (%d, %d) float64 device tensor"
                             % (self.dplan.n, ncol))
        return theta.detach().contiguous()

    def _write(self, theta):
        """Fields (and the Moffat values that follow beta) of theta into the device records."""
        import torch
        d = self.dplan
        aux = d._aux_dst.data_ptr() if (self.fit_beta and d._aux_dst is not None) else None
        rc = _lib.lib().gsmi_exp_moffat_fields(
            theta.data_ptr(), d.n, int(self.fit_beta), self._base.data_ptr(),
            self._cxcy.data_ptr(), 1. / self.scale, self._fwhm_to_rd.data_ptr(),
            d._dst_table.data_ptr(), aux, d._dev64.data_ptr(),
            torch.cuda.current_stream().cuda_stream)
        _lib.check(rc, "exp_moffat_fields")

    def _grad(self, theta, gsorted):
        import torch
        d = self.dplan
        out = torch.empty_like(theta)
        rc = _lib.lib().gsmi_exp_moffat_grad(
            theta.data_ptr(), d.n, int(self.fit_beta), self._fwhm_to_rd.data_ptr(),
            gsorted.data_ptr(), d._inv.data_ptr(), 1. / self.scale, out.data_ptr(),
            torch.cuda.current_stream().cuda_stream)
        _lib.check(rc, "exp_moffat_grad")
        return out


def _exp_moffat_function():
    import torch

    class ExpMoffatRender(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model):
            theta = model._check_theta(theta)
            model._write(theta)
            out = model.dplan.forward().clone()
            ctx.model = model
            ctx.save_for_backward(theta)
            return out

        @staticmethod
        def backward(ctx, gout):
            theta, = ctx.saved_tensors
            model = ctx.model
            model._write(theta)                  # another forward may have rewritten them
            return model._grad(theta, model.dplan.adjoint_sorted(gout)), None

    return ExpMoffatRender


_ExpMoffatRender = None


def moffat_aux(beta):
    """The Moffat record values that follow beta on the k^nu K_nu form (nu = beta - 1), as a
    (n, 5) tensor: the k = 0 value 2^(nu-1) Gamma(nu) = flux / knorm, and the order-only
    constants of Temme's series (moffat.temme_constants)."""
    import torch
    nu = beta - 1.
    at0 = torch.exp((nu - 1.) * math.log(2.) + torch.lgamma(nu))
    mu = nu - torch.floor(nu + 0.5)
    gampl = torch.exp(-torch.lgamma(1. + mu))
    gammi = torch.exp(-torch.lgamma(1. - mu))
    small = mu.abs() < 1e-4
    mus = torch.where(small, torch.ones_like(mu), mu)
    m2 = mu * mu
    gam1 = torch.where(small, -(0.5772156649015329 - 0.0420026350340952 * m2),
                       (gammi - gampl) / (2. * mus))
    gam2 = torch.where(small, 1. - 0.6558780715202538 * m2 + 0.1665386113822915 * m2 * m2,
                       (gammi + gampl) / 2.)
    fact = torch.where(small, 1. + (math.pi * mu) ** 2 / 6.,
                       math.pi * mus / torch.sin(math.pi * mus))
    return torch.stack([at0, fact * gam1, fact * gam2, 0.5 / gampl, 0.5 / gammi], dim=1)


def exp_moffat_fields(theta, base, cx, cy, scale, fwhm_to_rd, beta=None):
    """Fields of the Exponential x Moffat model as torch expressions of theta = (flux, hlr, g1,
    g2, dx, dy, psf_fwhm), following the lowering: the Exponential leaf has scale radius
    hlr / 1.6783469900166605 (galsim/exponential.py) and composite Jacobian diag(1/scale) S(g)
    with S(g) = [[1+g1, g2], [g2, 1-g1]] / sqrt(1-|g|^2) (galsim/shear.py:365-386); the Moffat
    leaf has rD = fwhm / (2 sqrt(2^(1/beta) - 1)) (galsim/moffat.py); the drawImage offset is
    the whole-result phase (ux, uy) = (dx, dy) + the centring shift (cx, cy).  With ``beta``
    (a tensor) the Moffat beta is a parameter too: knorm = 4 / (Gamma(beta-1) 2^beta) (unit PSF
    flux, the k^nu K_nu form), rD from it, and the shape field beta.  Fields the parameters do
    not reach come from ``base``."""
    import torch
    P = PER_LEAF
    flux, hlr, g1, g2, dx, dy, fw = theta.unbind(1)
    inv = 1. / scale
    den = torch.sqrt(1. - (g1 * g1 + g2 * g2))
    b = base
    if beta is None:
        m_flux, m_scale, m_shape = b[:, P], fw * fwhm_to_rd, b[:, P + 6]
    else:
        m_flux = 4. * torch.exp(-torch.lgamma(beta - 1.) - beta * math.log(2.))
        m_scale = fw / (2. * torch.sqrt(torch.pow(2., 1. / beta) - 1.))
        m_shape = beta
    cols = [flux, hlr / 1.6783469900166605,
            inv * (1. + g1) / den, inv * g2 / den, inv * g2 / den, inv * (1. - g1) / den,
            b[:, 6],
            m_flux, m_scale, b[:, P + 2], b[:, P + 3], b[:, P + 4], b[:, P + 5], m_shape]
    cols += [b[:, k] for k in range(2 * P, F_UX)]
    cols += [dx + cx, dy + cy, b[:, F_FS]]
    return torch.stack(cols, dim=1)
