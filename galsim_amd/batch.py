"""Batched rendering: thousands of stamps of uneven size in a handful of launches.

``draw_batch`` applies drawImage's per-stamp steps (offset convention, world->image transform,
pixel convolution, drawFFT_makeKImage sizing) to every profile, lowers them all into one device
buffer, bins the stamps by FFT size N and runs the two kernels once per bin on one stream.  The
result stays on the GPU as one flat tensor plus per-stamp offsets and shapes.
"""
import ctypes
import math

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
        p = Convolve(p, Pixel(scale=1.0, gsparams=prof.gsparams),
                     real_space=None if method == 'auto' else False, gsparams=prof.gsparams)
    b = BoundsI(1, nx, 1, ny)
    c = b.center
    return p, b.shift(PositionI(-c.x, -c.y)), flux_scale


# per-stamp status bits (BatchPlan.status); STAMP_NONFINITE (1) is the kernels' own
STAMP_BAD_PARAMS = 2        # a parameter or size was not finite: the stamp was not drawn
STAMP_FFT_TOO_LARGE = 4     # Nk > gsparams.maximum_fft_size (drawn; the reference warns)
STAMP_RERUN_FP64 = 8        # redrawn on the FP64 per-stamp path by rerun_flagged


def _plan_stamp(ctx, prof, nx, ny, scale, offset, use_true_center, method, precision,
                differentiable, moffat_generic):
    """One stamp's sizes (drawFFT_makeKImage, galsim/gsobject.py:1927-1973) and lowering into
    ctx.  Returns (record, N, warn) with warn set where Nk exceeds maximum_fft_size; raises
    ValueError for non-finite parameters or sizes."""
    p, b, _ = _image_profile(prof, nx, ny, scale, offset, use_true_center, method)
    # a non-finite size stops here (int() of NaN or inf); non-finite constants are found by
    # _nonfinite_stamps over the whole batch
    N, Nk, dk = render.plan_sizes(p, b)
    warn = Nk > p.gsparams.maximum_fft_size
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
        naxis += _prepare_differentiable(ctx, op0, nops, naxis, leaf0, moffat_generic)
    rec = (N, Nk, nx, ny, b.xmin, b.ymin, op0, nops, leaf0, naxis, 0, 0, 0, dk, 0, 0)
    return rec, N, warn


def _nonfinite_stamps(recs, ops, leaves):
    """Indices of the records whose program constants or leaf constants are not finite."""
    bad_op = ~(np.isfinite(ops['a']) & np.isfinite(ops['b']) & np.isfinite(ops['c']))
    bad_leaf = ~(np.isfinite(leaves['jac']).all(axis=1) & np.isfinite(leaves['scale']) &
                 np.isfinite(leaves['scale2']) & np.isfinite(leaves['flux']) &
                 np.isfinite(leaves['p']).all(axis=1))
    cop = np.concatenate([[0], np.cumsum(bad_op)])
    clf = np.concatenate([[0], np.cumsum(bad_leaf)])
    # each stamp's ops and leaves are contiguous blocks that do not interleave: a block runs to
    # the next block's start
    def ends(starts, total):
        o = np.argsort(starts, kind='stable')
        e = np.empty_like(starts)
        e[o] = np.append(starts[o][1:], total)
        return e
    o0 = recs['op_begin'].astype(np.int64)
    l0 = recs['leaf_begin'].astype(np.int64)
    o1 = ends(o0, len(ops))
    l1 = ends(l0, len(leaves))
    return np.nonzero((cop[o1] - cop[o0] > 0) | (clf[l1] - clf[l0] > 0))[0]


def _special_knorm2(kind):
    """knorm / flux of the closed-form Moffat kinds (src/SBMoffat.cpp:497-549)."""
    return {15: 1., 2: 1., 25: 1., 3: 0.5, 35: 1. / 3., 4: 0.125}[kind]


_RADIAL = (_lib.LEAF_EXPONENTIAL, _lib.LEAF_GAUSSIAN, _lib.LEAF_MOFFAT, _lib.LEAF_KOLMOGOROV,
           _lib.LEAF_SERSIC)


def _prepare_differentiable(ctx, op0, nops, naxis, leaf0, moffat_generic=False):
    """Make a lowered stamp's fields writable and differentiable: radial leaves of the program
    take the general-Jacobian path (so every jac entry is live), and the hoisted factors get a
    whole-result PHASE (ux = uy = 0, fs = 1, an exact no-op) if they have none.  With
    ``moffat_generic`` Moffat leaves use the k^nu K_nu form for every beta (the special-beta
    closed forms equal it at their beta, but only there), so beta can move.  Returns the number
    of ops appended.  Exponential clip rows (FP64) stay those of the lowered Jacobian."""
    from .moffat import temme_constants
    for code, arg, *_ in ctx.ops[op0:op0 + nops]:
        if code == _lib.OP_INV:
            raise NotImplementedError("a Deconvolution is not differentiable here")
        if code == _lib.OP_LEAF:
            rec = ctx.leaves[leaf0 + arg]
            if rec['type'] not in _RADIAL + (_lib.LEAF_BOX,):
                raise NotImplementedError("leaf type %d has no adjoint here" % rec['type'])
            if rec['type'] in _RADIAL:
                rec['separable'] = 0
            if moffat_generic and rec['type'] == _lib.LEAF_MOFFAT and rec['kind'] == 1:
                raise NotImplementedError("the beta of a truncated Moffat is not differentiable")
            if moffat_generic and rec['type'] == _lib.LEAF_MOFFAT and rec['kind'] != 0:
                beta = float(rec['p'][0])
                knorm2 = 4. / (math.gamma(beta - 1.) * 2. ** beta)
                # knorm = profile flux x knorm2: swap the closed form's knorm2 for the generic one
                rec['flux'] = float(rec['flux']) * knorm2 / _special_knorm2(int(rec['kind']))
                rec['kind'] = 0
                rec['p'][1] = 1. / knorm2
                rec['p'][2:6] = temme_constants(beta - 1.)
    axis = ctx.ops[op0 + nops:op0 + nops + naxis]
    if any(op[0] == _lib.OP_PHASE for op in axis):
        return 0
    ctx.emit(_lib.OP_PHASE, 0, 0., 0., 1.)
    return 1


def _program_shape(recs, ops, leaves):
    """The specialised program shape (_lib.STATIC_PROGRAMS id) every stamp of the batch has, or
    0 if they do not all share one."""
    found = None
    for r in recs:
        b, cnt, lb = int(r['op_begin']), int(r['op_count']), int(r['leaf_begin'])
        prog = tuple((int(o['code']), int(o['arg'])) for o in ops[b:b + cnt])
        shape = 0
        for sid, (sops, types) in _lib.STATIC_PROGRAMS.items():
            if prog == sops and all(int(leaves[lb + l]['type']) == t for l, t in enumerate(types)):
                shape = sid
                break
        if shape == 0 or (found is not None and shape != found):
            return 0
        found = shape
    return found or 0


FOLD_ITEM_CELLS = 16384       # default size of a fold work item (columns x N), in cells
LIST_LDS_LIMIT = 163840       # LDS one workgroup may allocate on gfx950 (160 KiB)


def lds_line(n):
    """Length of one transform line in the kernels' padded LDS layout (a slot after every 32)."""
    return n + n // 32
ROW_ITEM_CELLS = 1            # default size of a row work item (rows x N/2): one chunk
STAGE_LDS = 32768             # LDS per fold workgroup allowed for staging (5 workgroups per CU)


def work_lists(recs, groups, cols, rows, elem, bins=None, chunks=None, extra=0,
               row_chunks=None):
    """Fold and row work items (gsmi_work rows: stamp, first, count, per_group) of a batch whose
    stamp records ``recs`` are sorted into FFT-size ``groups`` (N, first, count, max_ny), for
    tiles ``cols`` / ``rows`` (dict N -> columns / rows per workgroup chunk; defaults 1024/N and
    1024/(N/2)), fold items largest FFT sizes first and row items in the reverse order, and the
    LDS bytes of a fold and of a row workgroup
    (elem: bytes per complex value; extra: bytes the fold kernel keeps beside its buffers).
    Every fold column 0..N/2-1 (column 0 carries kx = N/2 too) and every output row of every
    stamp is in exactly one item.

    A fold item of a size whose two transform buffers, N twiddles and 3 Nk/2 + 2 axis factors
    fit in STAGE_LDS spans ``chunks[N]`` chunks of ``cols[N]`` columns (default: about
    FOLD_ITEM_CELLS cells; the columns split evenly over the items), which share its set-up and
    its staged tables; the fold LDS then covers that staging.  Other sizes get one-chunk
    items.  A row item likewise spans ``row_chunks[N]`` chunks of ``rows[N]`` rows (default:
    about ROW_ITEM_CELLS cells) and stages the N twiddles, which the row LDS covers."""
    fold, rw = [], []
    groups = sorted(groups, key=lambda g: -g[0])
    if bins is not None:
        groups = [g for g in groups if g[0] in bins]
    chunks = chunks or {}
    row_chunks = row_chunks or {}
    lds_f = lds_r = 0
    for N, first, count, _ in groups:
        cpg = int(cols.get(N, BatchPlan.default_cols(N)))
        rpg = int(rows.get(N, BatchPlan.default_rows(N)))
        if cpg < 1 or rpg < 1:
            raise ValueError("tiles must hold at least one column and one row")
        W = N // 2               # fold columns: kx = 0 and kx = N/2 travel as one
        nk = int(recs['Nk'][first:first + count].max()) if count else 0
        bufs = 2 * cpg * lds_line(N) * elem + extra
        staged = bufs + (N + 3 * (nk // 2) + 2) * elem
        nchunk = -(-W // cpg)
        if staged <= STAGE_LDS:
            k = int(chunks.get(N, max(1, FOLD_ITEM_CELLS // (cpg * N))))
            if k < 1:
                raise ValueError("a fold item holds at least one chunk")
            lds_f = max(lds_f, staged)
        else:
            k = 1
            lds_f = max(lds_f, bufs)
        nitem = -(-nchunk // k)
        per_item = -(-nchunk // nitem) * cpg            # columns per item, a multiple of cpg
        rbufs = 2 * rpg * lds_line(N // 2) * elem
        lds_r = max(lds_r, rbufs + N * elem if rbufs + N * elem <= 65536 else rbufs)
        kr = int(row_chunks.get(N, max(1, ROW_ITEM_CELLS // (rpg * (N // 2)))))
        if kr < 1:
            raise ValueError("a row item holds at least one chunk")
        for s in range(first, first + count):
            for c0 in range(0, W, per_item):
                fold.append((s, c0, min(per_item, W - c0), cpg))
            ny = int(recs['ny'][s])
            nrc = -(-ny // rpg)
            rows_item = -(-nrc // -(-nrc // kr)) * rpg     # rows per item, a multiple of rpg
            for r0 in range(0, ny, rows_item):
                rw.append((s, r0, min(rows_item, ny - r0), rpg))
    if max(lds_f, lds_r) > LIST_LDS_LIMIT:
        raise _lib.GalSimAMDError("work-list tiles need %d bytes of LDS (limit %d)"
                                  % (max(lds_f, lds_r), LIST_LDS_LIMIT))
    # the row launch takes its items in the reverse of the fold's order: it starts on the stamps
    # the fold finished last, whose column-transform rows are still in the caches
    rw.reverse()
    return (np.array(fold, dtype=np.int32).reshape(-1, 4),
            np.array(rw, dtype=np.int32).reshape(-1, 4), lds_f, lds_r)


class HostPlan:
    """A batch lowered on the host: stamp records in input order (``keep`` marks the planned
    ones), the program, leaf, table and clip-row arrays of include/gsmi.h, per-stamp host
    status, the drawing arguments, and ``profile(i)``, the profile object of stamp i (for the
    FP64 re-run of flagged stamps)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def plan_host(profiles, nx, ny, scale, offsets=None, precision='f32', method='auto',
              use_true_center=True, differentiable=False, moffat_generic=False,
              on_error='flag'):
    """The per-object planner: drawImage's steps for each profile (image-coordinate profile,
    drawFFT_makeKImage sizes) and its lowering, stamp by stamp."""
    if precision not in ('f32', 'f64'):
        raise ValueError("precision must be 'f32' or 'f64'")
    if on_error not in ('flag', 'raise'):
        raise ValueError("on_error must be 'flag' or 'raise'")
    n = len(profiles)
    nx = np.broadcast_to(np.asarray(nx, dtype=np.int64), (n,)).copy()
    ny = np.broadcast_to(np.asarray(ny, dtype=np.int64), (n,)).copy()
    offsets = np.zeros((n, 2)) if offsets is None else np.asarray(offsets, dtype=float)
    ctx = render.Lowering(precision)
    recs = np.zeros(n, dtype=_lib.STAMP_DTYPE)
    # per-stamp status (input order): a stamp whose planning fails is left out of the launch
    # (drawn as zeros) and flagged instead of stopping the batch, unless on_error='raise'
    host_status = np.zeros(n, dtype=np.int32)
    keep = np.ones(n, dtype=bool)
    for i, prof in enumerate(profiles):
        nops0, nleaves0 = len(ctx.ops), len(ctx.leaves)
        try:
            rec, N, warn = _plan_stamp(ctx, prof, int(nx[i]), int(ny[i]), float(scale),
                                       offsets[i], use_true_center, method, precision,
                                       differentiable, moffat_generic)
        except (ValueError, OverflowError, ZeroDivisionError, _lib.GalSimAMDError):
            if on_error == 'raise':
                raise
            del ctx.ops[nops0:]
            del ctx.leaves[nleaves0:]
            host_status[i] |= STAMP_BAD_PARAMS
            keep[i] = False
            continue
        recs[i] = rec
        if warn:
            host_status[i] |= STAMP_FFT_TOO_LARGE
    # non-finite constants (a NaN or infinite flux, shear, offset...): flagged and dropped
    ops, leaves = ctx.ops_array(), ctx.leaves_array()
    cand = np.nonzero(keep)[0]
    for i in cand[_nonfinite_stamps(recs[cand], ops, leaves)]:
        if on_error == 'raise':
            raise ValueError("stamp %d has a non-finite parameter" % i)
        host_status[i] |= STAMP_BAD_PARAMS
        keep[i] = False
    profs = list(profiles)
    return HostPlan(n=n, precision=precision, recs=recs, keep=keep, host_status=host_status,
                    ops=ops, leaves=leaves,
                    tables=np.concatenate(ctx.tables) if ctx.tables else np.zeros(1),
                    tables32=ctx.tables32_array(),
                    clip=np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, dtype=np.int32),
                    nx=nx, ny=ny, scale=float(scale), offsets=offsets, method=method,
                    use_true_center=use_true_center, profile=lambda i: profs[i])


class BatchPlan:
    """A lowered batch, resident on the device.  ``run()`` launches the kernels on torch's
    current stream (no allocation, no host transfer: it can be captured in a HIP graph)."""

    def __init__(self, profiles, nx, ny, scale, offsets=None, precision='f32', method='auto',
                 use_true_center=True, differentiable=False, moffat_generic=False,
                 on_error='flag', lowered=None):
        """Plan ``profiles`` (drawn on nx x ny stamps at ``scale`` with ``offsets``) with the
        per-object planner, or take a ``HostPlan`` already lowered (``lowered``, e.g. from
        galsim_amd.config), and set the batch up on the device."""
        if lowered is None:
            lowered = plan_host(profiles, nx, ny, scale, offsets, precision, method,
                                use_true_center, differentiable, moffat_generic, on_error)
        self._setup(lowered)

    def _setup(self, hp):
        import torch
        n = hp.n
        precision = hp.precision
        self.precision = precision
        self.n = n
        self.lowered = hp
        self.host_status = hp.host_status.copy()
        keep = hp.keep
        Ns = hp.recs['N'].astype(np.int64)
        host_ops, host_leaves = hp.ops, hp.leaves
        self.dropped = np.nonzero(~keep)[0]
        # bin by N; stamp records sorted so each bin is contiguous (planned stamps only)
        kept = np.nonzero(keep)[0]
        order = kept[np.argsort(Ns[kept], kind='stable')]
        recs = hp.recs[order]
        self.order = order
        self.nrec = len(recs)
        def starts(sizes):
            # exclusive prefix sum (offsets of consecutive regions)
            return np.concatenate([[0], np.cumsum(sizes)[:-1]])[:len(sizes)].astype(np.int64)
        sizes = (recs['nx'].astype(np.int64) * recs['ny'].astype(np.int64))
        out_off = starts(sizes)
        recs['out_offset'] = out_off
        # scratch: every stamp's column-transform rows, then every stamp's axis factors, with
        # no region shared between stamps (a work-list launch runs all FFT sizes at once); the
        # columns are padded to a multiple of 16 values so each starts on a 128-byte line (the
        # row transform reads chunks of rows of every column)
        recs['col_stride'] = (recs['ny'] + 15) // 16 * 16
        rows = recs['col_stride'].astype(np.int64) * (recs['N'].astype(np.int64) // 2 + 1)
        axis = 3 * (recs['Nk'].astype(np.int64) // 2) + 2
        recs['scratch_offset'] = starts(rows)
        recs['axis_offset'] = int(rows.sum()) + starts(axis)
        scratch_need = int(rows.sum() + axis.sum())
        self.groups = []
        for N in np.unique(recs['N']):
            idx = np.nonzero(recs['N'] == N)[0]
            self.groups.append((int(N), int(idx[0]), len(idx), int(recs['ny'][idx].max())))
        pk = render._Packer()
        self._o_stamp = pk.add(recs)
        self._o_ops = pk.add(host_ops)
        self._o_leaves = pk.add(host_leaves)
        self._o_tab = pk.add(hp.tables)
        self._o_tab32 = pk.add(hp.tables32)
        self._o_clip = pk.add(hp.clip)
        self.recs = recs
        self.host_ops = host_ops
        self.host_leaves = host_leaves
        self.offsets = {'stamp': self._o_stamp, 'ops': self._o_ops, 'leaves': self._o_leaves,
                        'tables': self._o_tab, 'tables32': self._o_tab32, 'clip': self._o_clip}
        self.dev = torch.from_numpy(pk.host()).to('cuda')
        fdt = torch.float32 if precision == 'f32' else torch.float64
        self.scratch = torch.empty(2 * max(scratch_need, 1), dtype=fdt, device='cuda')
        self.out = torch.empty(int(sizes.sum()), dtype=fdt, device='cuda')
        # per input stamp: (offset, ny, nx) in the flat output (-1: a stamp left out)
        inv = np.full(n, -1, dtype=np.int64)
        inv[order] = np.arange(len(order))
        self.rec_of = inv
        self.stamp_offset = np.where(inv >= 0, out_off[np.maximum(inv, 0)] if len(order) else 0,
                                     -1)
        self.stamp_shape = np.stack([hp.ny, hp.nx], axis=1).astype(np.int64)
        self._rerun = np.zeros(n, dtype=np.int32)
        self._rerun_cleared = np.zeros(n, dtype=bool)
        self.nops = len(host_ops)
        self.nleaves = len(host_leaves)
        self.program_shape = _program_shape(recs, self.host_ops, self.host_leaves)
        self._sched = None
        self._graph = None
        self.tuning = None

    # -- work-list schedule ----------------------------------------------------------------------
    @staticmethod
    def default_cols(N):
        """Columns per fold workgroup of the per-size kernels: 1024 / N (one column from 1024)."""
        return max(1, 1024 // N)

    @staticmethod
    def default_rows(N):
        """Rows per row-transform workgroup of the per-size kernels: 1024 / (N/2)."""
        return max(1, 1024 // (N // 2))

    def _elem(self):
        return 8 if self.precision == 'f32' else 16

    def _fold_extra(self):
        """LDS bytes the specialised fold kernel of the batch's program shape keeps beside its
        buffers: 4 decoded leaves of 20 values, and in FP32 the cubic coefficients of the
        largest Sersic table (16 bytes per knot interval)."""
        if not self.program_shape:
            return 0
        extra = 4 * 20 * (4 if self.precision == 'f32' else 8)
        if self.precision == 'f32':
            sel = self.host_leaves['type'] == _lib.LEAF_SERSIC
            if sel.any():
                extra += 16 * int(self.host_leaves['ntab'][sel].max() - 1)
        return extra

    def _work_lists(self, cols, rows, bins=None, chunks=None):
        """work_lists with the specialised kernel's LDS extra, or, if a tile leaves no room for
        it, without (the batch then runs the interpreter kernel).  Returns the lists, the LDS
        sizes and the program shape to launch."""
        extra = self._fold_extra()
        try:
            return work_lists(self.recs, self.groups, cols, rows, self._elem(), bins, chunks,
                              extra) + (self.program_shape,)
        except _lib.GalSimAMDError:
            if not extra:
                raise
        return work_lists(self.recs, self.groups, cols, rows, self._elem(), bins, chunks,
                          0) + (0,)

    def make_schedule(self, cols=None, rows=None, fold_grid=None, rows_grid=None, bins=None,
                      chunks=None):
        """(device item tensor, gsmi_schedule) for tiles ``cols``/``rows`` (dict N -> columns or
        rows per workgroup chunk; missing sizes take the defaults), fold chunks per item
        (``chunks``, dict N -> count) and grids (None: one workgroup per item; smaller:
        persistent workgroups).  ``bins`` restricts it to some FFT sizes."""
        import torch
        fold, rw, lds_f, lds_r, program = self._work_lists(cols or {}, rows or {}, bins,
                                                           chunks)
        items = torch.from_numpy(np.ascontiguousarray(np.concatenate([fold, rw]))).cuda()
        sc = _lib.Schedule()
        base = items.data_ptr()
        sc.fold = base
        sc.rows = base + len(fold) * 16
        sc.nfold, sc.nrows = len(fold), len(rw)
        sc.fold_grid = len(fold) if fold_grid is None else int(min(fold_grid, len(fold)))
        sc.rows_grid = len(rw) if rows_grid is None else int(min(rows_grid, len(rw)))
        sc.fold_grid, sc.rows_grid = max(sc.fold_grid, 1), max(sc.rows_grid, 1)
        sc.fold_lds = lds_f
        sc.rows_lds = lds_r
        sc.program = program
        return items, sc

    def use_schedule(self, cols=None, rows=None, fold_grid=None, rows_grid=None, chunks=None):
        """Run every FFT size in one fold and one row launch from now on."""
        self._sched = self.make_schedule(cols, rows, fold_grid, rows_grid, chunks=chunks)
        self._graph = None

    def _launch_list(self, sched, part, s):
        lib = _lib.lib()
        base = self.dev.data_ptr()
        sc = ctypes.addressof(sched[1])
        args_fold = (base + self._o_stamp, self.nrec, sc, base + self._o_ops,
                     base + self._o_leaves,
                     base + self._o_tab)
        if part in ('fold', 'all'):
            if self.precision == 'f32':
                rc = lib.gsmi_fold_list_f32(*args_fold, base + self._o_tab32,
                                            self.scratch.data_ptr(), s)
            else:
                rc = lib.gsmi_fold_list_f64(*args_fold, base + self._o_clip,
                                            self.scratch.data_ptr(), s)
            _lib.check(rc, "fold work list")
        if part in ('rows', 'all'):
            f = lib.gsmi_rows_list_f32 if self.precision == 'f32' else lib.gsmi_rows_list_f64
            rc = f(base + self._o_stamp, self.nrec, sc, self.scratch.data_ptr(),
                   self.out.data_ptr(), s)
            _lib.check(rc, "row work list")

    # -- tuning and graphs -----------------------------------------------------------------------
    def _time(self, fn, reps):
        import torch
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    def autotune(self, reps=5, factors=(0.25, 0.5, 1, 2, 4), grids=(None, 1024, 2048, 4096),
                 item_cells=(1, 4096, FOLD_ITEM_CELLS, 65536)):
        """Pick, by timing on this batch, the columns per fold chunk and the rows per row
        workgroup of every FFT size (powers of two around the defaults, within the LDS budget),
        then the size of a fold item (``item_cells``: cells per item, 1 = one chunk) and the fold
        and row grid sizes (one workgroup per item, or persistent) on the whole launch, and
        switch run() to the tuned work lists.  Returns the choice and its timings (ms)."""
        import torch
        s = torch.cuda.current_stream().cuda_stream
        cols, rows, log = {}, {}, {}
        elem = self._elem()
        for N, first, count, _ in self.groups:
            W = N // 2               # fold columns (kx = 0 and kx = N/2 packed as one)
            best = None
            # powers of two around the default, and widths that split the W columns into
            # equal chunks
            d = self.default_cols(N)
            cands = {int(d * f) for f in factors}
            cands |= {-(-W // m) for m in range(max(1, W // (2 * d)), 2 * (-(-W // d)) + 1)}
            for c in sorted(cands):
                # the launch's LDS is the largest size's: keep tiles within the staging budget
                if c < 1 or c > W or 2 * c * lds_line(N) * elem + self._fold_extra() > \
                        (LIST_LDS_LIMIT if c == 1 else STAGE_LDS):
                    continue
                sched = self.make_schedule({N: c}, {}, bins=(N,))
                t = self._time(lambda: self._launch_list(sched, 'fold', s), reps)
                log[('cols', N, c)] = t
                if best is None or t < best[0]:
                    best = (t, c)
            cols[N] = best[1]
            best = None
            for f in factors:
                r = int(self.default_rows(N) * f)
                if r < 1 or r > int(self.recs['ny'][first:first + count].max()) * 2 or \
                        2 * r * lds_line(N // 2) * elem > LIST_LDS_LIMIT:
                    continue
                sched = self.make_schedule({N: cols[N]}, {N: r}, bins=(N,))
                self._launch_list(sched, 'fold', s)
                t = self._time(lambda: self._launch_list(sched, 'rows', s), reps)
                log[('rows', N, r)] = t
                if best is None or t < best[0]:
                    best = (t, r)
            rows[N] = best[1]
        # fold item size (cells per item, the same for every FFT size) x grid, on the whole
        # launch; the best few combinations are then timed again, interleaved over several
        # rounds, and the one with the lowest median wins (single short timings of launches
        # this size vary by a few percent from run to run)
        combos = []
        for cells in item_cells:
            ck = {N: max(1, cells // (cols[N] * N)) for N, _, _, _ in self.groups}
            for g in grids:
                sched = self.make_schedule(cols, rows, g, g, chunks=ck)
                t = self._time(lambda: self._launch_list(sched, 'all', s), reps)
                log[('item_cells', cells, 'grid', g)] = t
                combos.append((t, cells, g, ck))
        combos.sort(key=lambda c: c[0])
        finalists = combos[:4]
        scheds = [self.make_schedule(cols, rows, g, g, chunks=ck) for _, _, g, ck in finalists]
        rounds = [[] for _ in finalists]
        for _ in range(3):
            for k, sched in enumerate(scheds):
                rounds[k].append(self._time(lambda: self._launch_list(sched, 'all', s),
                                            2 * reps))
        med = [float(np.median(r)) for r in rounds]
        k = int(np.argmin(med))
        _, cells, g, chunks = finalists[k]
        best = (med[k], g)
        log[('final', cells, g)] = med[k]
        self.use_schedule(cols, rows, best[1], best[1], chunks)
        self.tuning = {'cols': cols, 'rows': rows, 'chunks': chunks, 'grid': best[1],
                       'ms': best[0],
                       'timings': {str(k): v for k, v in log.items()}}
        return self.tuning

    def capture(self):
        """Capture one run() in a HIP graph (torch.cuda.CUDAGraph); replay() then launches the
        whole batch with one graph launch."""
        import torch
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize()
        self._graph = g
        return g

    def replay(self):
        if self._graph is None:
            raise _lib.GalSimAMDError("capture() the batch first")
        self._reset_reruns()
        self._graph.replay()
        return self.out

    def _reset_reruns(self):
        # a new run overwrites the stamps rerun_flagged redrew
        self._rerun[:] = 0
        self._rerun_cleared[:] = False

    def run(self, stream=None, events=None):
        """Launch the batch: with a work-list schedule (use_schedule / autotune), the axis
        prepass, one fold and one row launch for every FFT size; otherwise each FFT size's
        kernels in turn.  ``events``, if given, is a callable ``events(kind, N)`` returning a
        torch.cuda.Event to record before each kernel ('fold' / 'rows', N = 0 for a work list)
        and after the last ('end', 0)."""
        import torch
        lib = _lib.lib()
        self._reset_reruns()
        base = self.dev.data_ptr()
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        sc = self.scratch.data_ptr()
        if self._sched is not None:
            if events:
                events('fold', 0).record()
            self._launch_list(self._sched, 'fold', s)
            if events:
                events('rows', 0).record()
            self._launch_list(self._sched, 'rows', s)
            if events:
                events('end', 0).record()
            return self.out
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
        """Stamp i (input order) as a (ny, nx) device tensor view (zeros for a stamp left out of
        the launch, see ``status``)."""
        ny, nx = (int(v) for v in self.stamp_shape[i])
        off = int(self.stamp_offset[i])
        if off < 0:
            import torch
            return torch.zeros((ny, nx), dtype=self.out.dtype, device='cuda')
        return self.out[off:off + ny * nx].view(ny, nx)

    def status(self):
        """Per-stamp status (input order), after a run: STAMP_BAD_PARAMS for a stamp whose
        parameters or sizes were not finite (not drawn), STAMP_FFT_TOO_LARGE where Nk exceeds
        gsparams.maximum_fft_size (drawn, as the reference warns at galsim/gsobject.py:1965),
        STAMP_NONFINITE for a stamp with a NaN or infinite pixel (the kernels' own check),
        STAMP_RERUN_FP64 for one redrawn by ``rerun_flagged``.  Reads the device records."""
        st = self.host_status.copy()
        if self.nrec:
            import torch
            torch.cuda.synchronize()
            d = self.dev[self._o_stamp:self._o_stamp + self.nrec * _lib.STAMP_DTYPE.itemsize]
            recs = d.cpu().numpy().view(_lib.STAMP_DTYPE)
            st[self.order] |= recs['status'].astype(np.int32)
        st |= self._rerun
        st[self._rerun_cleared] &= ~_lib.STAMP_NONFINITE
        return st

    def rerun_flagged(self):
        """Redraw every stamp whose pixels came out non-finite on the FP64 per-stamp path (an
        FP32 intermediate can overflow where the FP64 one does not) and copy it into the
        output.  Returns the input indices redrawn; their status gains STAMP_RERUN_FP64, and
        keeps STAMP_NONFINITE only if the FP64 pixels are not finite either."""
        import torch
        st = self.status()
        redo = np.nonzero((st & _lib.STAMP_NONFINITE) != 0)[0]
        hp = self.lowered
        for i in redo:
            im = hp.profile(i).drawImage(nx=int(hp.nx[i]), ny=int(hp.ny[i]), scale=hp.scale,
                                         offset=tuple(hp.offsets[i]), method=hp.method,
                                         use_true_center=hp.use_true_center, dtype=np.float64,
                                         precision='f64')
            a = torch.from_numpy(np.ascontiguousarray(im.array)).to('cuda', self.out.dtype)
            self.stamp(i).copy_(a)
            self._rerun[i] |= STAMP_RERUN_FP64
            if torch.isfinite(a).all():
                self._rerun_cleared[i] = True
        return redo

    def fft_sizes(self):
        return {N: count for N, _, count, _ in self.groups}


def draw_batch(profiles, nx, ny, scale, offsets=None, precision='f32', method='auto',
               use_true_center=True):
    """Render a list of profiles to stamps of shape (ny[i], nx[i]); returns the BatchPlan whose
    ``out`` holds the pixels (``stamp(i)`` gives one stamp)."""
    plan = BatchPlan(profiles, nx, ny, scale, offsets, precision, method, use_true_center)
    plan.run()
    return plan
