"""numpy emulation of the device k-space program (the leaf formulas and op semantics of
galsim_amd/csrc/gsmi_kernels.hip), used to check the host lowering on the CPU."""
import numpy as np
import scipy.special as sps

from galsim_amd import _lib
from galsim_amd import render


def _leaf(L, tables, clip, kx, ky, i, j, Nk):
    if L['separable']:
        kxp = L['jac'][0] * kx
        kyp = L['jac'][3] * ky
    else:
        kxp = L['jac'][0] * kx + L['jac'][2] * ky
        kyp = L['jac'][1] * kx + L['jac'][3] * ky
    X = kxp * L['scale']
    Y = kyp * L['scale']
    flux = L['flux']
    t = L['type']
    ksq = X * X + Y * Y
    if t == _lib.LEAF_EXPONENTIAL:
        v = flux / ((1 + ksq) * np.sqrt(1 + ksq))
        if L['cliprows'] >= 0:
            r = clip[L['cliprows']:L['cliprows'] + 2 * (Nk + 1)].reshape(-1, 2)[j + Nk // 2]
            keep = (i >= r[..., 0]) & (i < r[..., 1])
        else:
            keep = ksq <= L['p'][0]
        return np.where(keep, v, 0.)
    if t == _lib.LEAF_GAUSSIAN:
        if L['separable']:
            return flux * np.exp(-0.5 * X * X) * np.exp(-0.5 * Y * Y)
        v = np.where(ksq < L['p'][1], flux * (1 - 0.5 * ksq * (1 - 0.25 * ksq)),
                     flux * np.exp(-0.5 * ksq))
        return np.where(ksq > L['p'][0], 0., v)
    if t == _lib.LEAF_MOFFAT:
        k = np.sqrt(ksq)
        kind = L['kind']
        with np.errstate(invalid='ignore', divide='ignore'):
            if kind == 15:
                return flux * np.exp(-k)
            if kind == 25:
                return flux * np.exp(-k) * (1 + k)
            if kind == 35:
                return flux * np.exp(-k) * (3 + (3 + k) * k)
            if kind == 2:
                return np.where(ksq == 0, flux, flux * sps.kv(1, k) * k)
            if kind == 3:
                return np.where(ksq == 0, 2 * flux, flux * sps.kv(2, k) * ksq)
            if kind == 4:
                return np.where(ksq == 0, 8 * flux, flux * sps.kv(3, k) * k * ksq)
            nu = L['p'][0] - 1
            return np.where(ksq == 0, flux * L['p'][1], flux * sps.kv(nu, k) * k ** nu)
    if t == _lib.LEAF_KOLMOGOROV:
        with np.errstate(divide='ignore'):
            return flux * np.exp(-np.exp(5. / 6. * np.log(ksq)))
    if t == _lib.LEAF_SERSIC:
        n = L['ntab']
        tab = tables[L['table']:L['table'] + 3 * n]
        args, vals, y2 = tab[:n], tab[n:2 * n], tab[2 * n:]
        p = L['p']
        with np.errstate(divide='ignore', invalid='ignore'):
            lk = 0.5 * np.log(ksq)
            a0 = args[0]
            da = (args[-1] - a0) / (n - 1)
            ii = np.clip(np.ceil((lk - a0) / da), 1, n - 1).astype(np.int64)
            ii = np.where(np.isfinite(lk), ii, 1)
            h = args[ii] - args[ii - 1]
            aa = args[ii] - lk
            bb = h - aa
            spl = (aa * vals[ii - 1] + bb * vals[ii] -
                   (1. / 6.) * aa * bb * ((aa + h) * y2[ii - 1] + (bb + h) * y2[ii])) / h
            spl = np.where((lk < a0) | (lk > args[-1]), 0., spl)
            mid = flux * spl / ksq
            hi = flux * (p[2] + p[3] / np.sqrt(ksq)) / ksq
            lo = flux * (1 + ksq * (p[4] + ksq * p[5]))
        return np.where(ksq >= p[1], hi, np.where(ksq < p[0], lo, mid))
    if t == _lib.LEAF_BOX:
        Xb = kxp * L['scale']
        Yb = kyp * L['scale2']
        return flux * np.sinc(Xb) * np.sinc(Yb)
    raise ValueError(t)


def eval_program(ctx, i, j, dk, Nk):
    """Evaluate a render.Lowering's program at integer k-grid points (i, j) (numpy arrays)."""
    tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
    clip = np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, np.int32)
    kx = i * dk
    ky = j * dk
    st = []
    for code, arg, a, b, c in ctx.ops:
        if code == _lib.OP_LEAF:
            st.append(_leaf(ctx.leaves[arg], tables, clip, kx, ky, i, j, Nk).astype(complex))
        elif code == _lib.OP_MUL:
            y = st.pop()
            st[-1] = st[-1] * y
        elif code == _lib.OP_ADD:
            y = st.pop()
            st[-1] = st[-1] + y
        elif code == _lib.OP_SCALE:
            st[-1] = st[-1] * a
        elif code == _lib.OP_PHASE:
            arg_ = a * kx + b * ky
            st[-1] = st[-1] * (c * np.exp(-1j * arg_))
    assert len(st) == 1
    return st[0]


def eval_axis(ctx, axis_ops, i, j, dk):
    """The hoisted separable factors X(kx) Y(ky) (axis_factor in the kernel)."""
    kx = i * dk
    ky = j * dk
    f = np.ones(np.broadcast(i, j).shape, dtype=complex)
    for code, arg, a, b, c in axis_ops:
        if code == _lib.OP_LEAF:
            L = ctx.leaves[arg]
            f = f * (L['flux'] * np.sinc(L['jac'][0] * kx * L['scale']) *
                     np.sinc(L['jac'][3] * ky * L['scale2']))
        elif code == _lib.OP_SCALE:
            f = f * a
        elif code == _lib.OP_PHASE:
            f = f * (c * np.exp(-1j * a * kx) * np.exp(-1j * b * ky))
    return f


def emulate_draw(prof, bounds, hoist=False):
    """Plan and lower ``prof`` exactly as render.draw_fft does, evaluate the program on the
    stored k-image and finish with the oracle's wrap/irfft/crop."""
    from oracle import draw as od
    N, Nk, dk = render.plan_sizes(prof, bounds)
    ctx = render.Lowering('f64', Nk, dk)
    prof._lower(ctx, None)
    ii, jj = np.meshgrid(np.arange(0, Nk // 2 + 1), np.arange(-Nk // 2, Nk // 2 + 1))
    if hoist:
        nops, naxis = render.hoist_axis_factors(ctx, 0)
        axis_ops = ctx.ops[nops:]
        ctx.ops = ctx.ops[:nops]
        karr = eval_program(ctx, ii, jj, dk, Nk) * eval_axis(ctx, axis_ops, ii, jj, dk)
    else:
        karr = eval_program(ctx, ii, jj, dk, Nk)
    kw = od.wrap_hermx(karr, Nk, N)
    real = od.irfft_shifted(kw, N)
    b = bounds
    return real[b.ymin + N // 2: b.ymax + N // 2 + 1, b.xmin + N // 2: b.xmax + N // 2 + 1]
