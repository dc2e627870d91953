"""Batched rendering: thousands of stamps of uneven size in a handful of launches.

``draw_batch`` applies drawImage's per-stamp steps (offset convention, world->image transform,
pixel convolution, drawFFT_makeKImage sizing) to every profile, lowers them all into one device
buffer, bins the stamps by FFT size N and runs the two kernels once per bin on one stream.  The
result stays on the GPU as one flat tensor plus per-stamp offsets and shapes.
"""
import numpy as np

from . import _lib
from . import render
from .box import Pixel
from .convolve import Convolve
from .gsobject import GSObject
from .position import BoundsI, PositionD, PositionI
from .transform import _Transform


def _image_profile(prof, nx, ny, scale, offset, use_true_center, method):
    """The image-coordinate profile and centred bounds drawImage would draw
    (galsim/gsobject.py:1720-1777)."""
    off = GSObject._adjust_offset(BoundsI(1, nx, 1, ny), PositionD(offset), None,
                                  use_true_center)
    jac = None if scale == 1. else np.array([[1. / scale, 0.], [0., 1. / scale]])
    flux_scale = 1. / (scale * scale) if method == 'sb' else 1.
    p = _Transform(prof, jac, (off.x, off.y), scale ** 2 * flux_scale)
    if method in ('auto', 'fft'):
        p = Convolve(p, Pixel(scale=1.0, gsparams=prof.gsparams), real_space=False,
                     gsparams=prof.gsparams)
    b = BoundsI(1, nx, 1, ny)
    c = b.center
    return p, b.shift(PositionI(-c.x, -c.y)), flux_scale


_RADIAL = (_lib.LEAF_EXPONENTIAL, _lib.LEAF_GAUSSIAN, _lib.LEAF_MOFFAT, _lib.LEAF_KOLMOGOROV,
           _lib.LEAF_SERSIC)


def _prepare_differentiable(ctx, op0, nops, naxis, leaf0):
    """Make a lowered stamp's fields writable and differentiable: radial leaves of the program
    take the general-Jacobian path (so every jac entry is live), and the hoisted factors get a
    whole-result PHASE (ux = uy = 0, fs = 1, an exact no-op) if they have none.  Returns the
    number of ops appended.  Exponential clip rows (FP64) stay those of the lowered Jacobian."""
    for code, arg, *_ in ctx.ops[op0:op0 + nops]:
        if code == _lib.OP_LEAF:
            rec = ctx.leaves[leaf0 + arg]
            if rec['type'] in _RADIAL:
                rec['separable'] = 0
    axis = ctx.ops[op0 + nops:op0 + nops + naxis]
    if any(op[0] == _lib.OP_PHASE for op in axis):
        return 0
    ctx.emit(_lib.OP_PHASE, 0, 0., 0., 1.)
    return 1


class BatchPlan:
    """A lowered batch, resident on the device.  ``run()`` launches the kernels on torch's
    current stream (no allocation, no host transfer: it can be captured in a HIP graph)."""

    def __init__(self, profiles, nx, ny, scale, offsets=None, precision='f32', method='auto',
                 use_true_center=True, differentiable=False):
        import torch
        if precision not in ('f32', 'f64'):
            raise ValueError("precision must be 'f32' or 'f64'")
        n = len(profiles)
        nx = np.broadcast_to(np.asarray(nx, dtype=np.int64), (n,))
        ny = np.broadcast_to(np.asarray(ny, dtype=np.int64), (n,))
        if offsets is None:
            offsets = np.zeros((n, 2))
        self.precision = precision
        self.n = n
        ctx = render.Lowering(precision)
        recs = np.zeros(n, dtype=_lib.STAMP_DTYPE)
        Ns = np.zeros(n, dtype=np.int64)
        for i, prof in enumerate(profiles):
            p, b, _ = _image_profile(prof, int(nx[i]), int(ny[i]), float(scale), offsets[i],
                                     use_true_center, method)
            N, Nk, dk = render.plan_sizes(p, b)
            ctx.Nk, ctx.dk = Nk, dk
            op0, leaf0 = len(ctx.ops), len(ctx.leaves)
            p._lower(ctx, None)
            nops, naxis = render.hoist_axis_factors(ctx, op0)
            for k in range(op0, len(ctx.ops)):
                code, arg, a, bb, c = ctx.ops[k]
                if code == _lib.OP_LEAF:
                    ctx.ops[k] = (code, arg - leaf0, a, bb, c)
            render.check_stack(ctx.ops, op0, nops, precision)
            if differentiable:
                naxis += _prepare_differentiable(ctx, op0, nops, naxis, leaf0)
            recs[i] = (N, Nk, nx[i], ny[i], b.xmin, b.ymin, op0, nops, leaf0, naxis,
                       0, 0, 0, dk)
            Ns[i] = N
        # bin by N; stamp records sorted so each bin is contiguous
        order = np.argsort(Ns, kind='stable')
        recs = recs[order]
        self.order = order
        out_off = np.zeros(n, dtype=np.int64)
        sizes = (recs['nx'].astype(np.int64) * recs['ny'].astype(np.int64))
        out_off[1:] = np.cumsum(sizes)[:-1]
        recs['out_offset'] = out_off
        self.groups = []
        scratch_need = 0
        for N in np.unique(recs['N']):
            idx = np.nonzero(recs['N'] == N)[0]
            w = N // 2 + 1
            rows = recs['ny'][idx].astype(np.int64) * w
            axis = 3 * (recs['Nk'][idx].astype(np.int64) // 2) + 2
            s_off = np.zeros(len(idx), dtype=np.int64)
            s_off[1:] = np.cumsum(rows)[:-1]
            a_off = int(rows.sum()) + np.concatenate([[0], np.cumsum(axis)[:-1]])
            recs['scratch_offset'][idx] = s_off
            recs['axis_offset'][idx] = a_off
            scratch_need = max(scratch_need, int(rows.sum() + axis.sum()))
            self.groups.append((int(N), int(idx[0]), len(idx), int(recs['ny'][idx].max())))
        pk = render._Packer()
        self._o_stamp = pk.add(recs)
        self._o_ops = pk.add(ctx.ops_array())
        self._o_leaves = pk.add(ctx.leaves_array())
        tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
        self._o_tab = pk.add(tables)
        self._o_tab32 = pk.add(tables.astype(np.float32))
        clip = np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, dtype=np.int32)
        self._o_clip = pk.add(clip)
        self.recs = recs
        self.host_ops = ctx.ops_array()
        self.host_leaves = ctx.leaves_array()
        self.offsets = {'stamp': self._o_stamp, 'ops': self._o_ops, 'leaves': self._o_leaves,
                        'tables': self._o_tab, 'tables32': self._o_tab32, 'clip': self._o_clip}
        self.dev = torch.from_numpy(pk.host()).to('cuda')
        fdt = torch.float32 if precision == 'f32' else torch.float64
        self.scratch = torch.empty(2 * max(scratch_need, 1), dtype=fdt, device='cuda')
        self.out = torch.empty(int(sizes.sum()), dtype=fdt, device='cuda')
        # per input stamp: (offset, ny, nx) in the flat output
        inv = np.empty(n, dtype=np.int64)
        inv[order] = np.arange(n)
        self.stamp_offset = out_off[inv]
        self.stamp_shape = np.stack([recs['ny'][inv], recs['nx'][inv]], axis=1)
        self.nops = len(ctx.ops)
        self.nleaves = len(ctx.leaves)

    def run(self, stream=None, events=None):
        """Launch every bin's two kernels.  ``events``, if given, is a callable
        ``events(kind, N)`` returning a torch.cuda.Event to record before each kernel
        ('fold' / 'rows') and after the last ('end', 0)."""
        import torch
        lib = _lib.lib()
        base = self.dev.data_ptr()
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        sc = self.scratch.data_ptr()
        for N, first, count, max_ny in self.groups:
            st = base + self._o_stamp + first * _lib.STAMP_DTYPE.itemsize
            if events:
                events('fold', N).record()
            if self.precision == 'f32':
                rc = lib.gsmi_fold_colfft_f32(st, count, N, base + self._o_ops,
                                              base + self._o_leaves, base + self._o_tab,
                                              base + self._o_tab32, sc, s)
            else:
                rc = lib.gsmi_fold_colfft_f64(st, count, N, base + self._o_ops,
                                              base + self._o_leaves, base + self._o_tab,
                                              base + self._o_clip, sc, s)
            _lib.check(rc, "batched fold/column FFT (N=%d)" % N)
            if events:
                events('rows', N).record()
            if self.precision == 'f32':
                rc = lib.gsmi_row_c2r_f32(st, count, N, max_ny, sc, self.out.data_ptr(), s)
            else:
                rc = lib.gsmi_row_c2r_f64(st, count, N, max_ny, sc, self.out.data_ptr(), s)
            _lib.check(rc, "batched row FFT (N=%d)" % N)
        if events:
            events('end', 0).record()
        return self.out

    def stamp(self, i):
        """Stamp i (input order) as a (ny, nx) device tensor view."""
        off = int(self.stamp_offset[i])
        ny, nx = (int(v) for v in self.stamp_shape[i])
        return self.out[off:off + ny * nx].view(ny, nx)

    def fft_sizes(self):
        return {N: count for N, _, count, _ in self.groups}


def draw_batch(profiles, nx, ny, scale, offsets=None, precision='f32', method='auto',
               use_true_center=True):
    """Render a list of profiles to stamps of shape (ny[i], nx[i]); returns the BatchPlan whose
    ``out`` holds the pixels (``stamp(i)`` gives one stamp)."""
    plan = BatchPlan(profiles, nx, ny, scale, offsets, precision, method, use_true_center)
    plan.run()
    return plan
