"""Host planning and device execution of the FFT drawing path.

``draw_fft`` restates drawFFT_makeKImage's size planning (galsim/gsobject.py:1927-1973), lowers
the profile tree into the k-space program of include/gsmi.h, packs the stamp record, program,
leaves, Sersic tables and Exponential clip rows into one host buffer, copies it to the device
in one transfer, and runs the two kernels of gsmi_draw_fft_f64/f32 on torch's current stream.
"""
import math
import warnings

import numpy as np

from . import _lib

_ALIGN = 256


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.GalSimAMDError("galsim_amd draws on the GPU; no HIP device is available")
    return torch


def spline_coefficients(args, vals, y2):
    """Per-interval (c0, c1, c2, c3), FP64, of the natural cubic spline through (args, vals) with
    second derivatives y2: on interval i (args[i]..args[i+1], width h) the spline is
    c0 + s (c1 + s (c2 + s c3)), s = (x - args[i]) / h -- Table's spline (src/Table.cpp:560-577)
    written as a polynomial in s."""
    a = np.asarray(args, float)
    v = np.asarray(vals, float)
    y = np.asarray(y2, float)
    hh = np.diff(a) ** 2
    c = np.empty((len(a) - 1, 4))
    c[:, 0] = v[:-1]
    c[:, 1] = (v[1:] - v[:-1]) - hh * (2. * y[:-1] + y[1:]) / 6.
    c[:, 2] = 0.5 * hh * y[:-1]
    c[:, 3] = hh * (y[1:] - y[:-1]) / 6.
    return c


class LeafRecord:
    """A gsmi_leaf under construction: the same field names (rec['scale'] = ..., rec['p'][k] =
    ...), held in Python values until ``Lowering.leaves_array`` packs the batch column by
    column (a numpy structured scalar per leaf costs more than the rest of its lowering)."""
    __slots__ = ('d',)
    _names = ('type', 'separable', 'kind', 'table', 'ntab', 'cliprows', 'jac', 'scale', 'scale2',
              'flux', 'p')

    def __init__(self, d=None):
        self.d = d if d is not None else {
            'type': 0, 'separable': 0, 'kind': 0, 'table': 0, 'ntab': 0, 'cliprows': -1,
            'jac': [1., 0., 0., 1.], 'scale': 0., 'scale2': 0., 'flux': 0., 'p': [0.] * 6}

    def __getitem__(self, k):
        return self.d[k]

    def __setitem__(self, k, v):
        if k not in self.d:
            raise KeyError(k)
        if k in ('jac', 'p'):
            v = [float(x) for x in v]
            if len(v) != len(self.d[k]):
                raise ValueError("%s takes %d values" % (k, len(self.d[k])))
        self.d[k] = v

    def copy(self):
        d = dict(self.d)
        d['jac'] = list(d['jac'])
        d['p'] = list(d['p'])
        return LeafRecord(d)


class Lowering:
    """Accumulates the k-space program of one or more stamps."""

    def __init__(self, precision='f64', Nk=None, dk=None):
        self.precision = precision
        self.Nk = Nk
        self.dk = dk
        self.ops = []
        self.leaves = []
        self.tables = []
        self.tab_off = {}
        self.ntab = 0
        self.clip = []
        self.nclip = 0
        self.coefs = []
        self.coef_off = {}
        self.ncoef = 0
        self.kvalue_mode = False     # lowering for kValue: no k-image cuts
        self.kbounds = None          # (xmin, ymin, columns, rows) of a k-image not drawFFT's
        self.owners = []

    def emit(self, code, arg=0, a=0., b=0., c=0.):
        self.ops.append((code, arg, a, b, c))

    def add_leaf(self, rec):
        self.leaves.append(rec.copy())
        return len(self.leaves) - 1

    def add_table(self, key, args, vals, y2):
        if key not in self.tab_off:
            self.tab_off[key] = self.ntab
            self.tables.append(np.concatenate([args, vals, y2]))
            self.ntab += 3 * len(args)
        return self.tab_off[key]

    def add_block(self, key, arr, owner=None):
        """Append a block of doubles to the table buffer, once per ``key`` (None: always);
        returns its offset.  ``owner`` is kept alive with the lowering (keys built from id())."""
        if key is not None and key in self.tab_off:
            return self.tab_off[key]
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        off = self.ntab
        self.tables.append(a)
        self.ntab += a.size
        if key is not None:
            self.tab_off[key] = off
        if owner is not None:
            self.owners.append(owner)
        return off

    def add_spline_coefs(self, key, args, vals, y2):
        """The natural cubic spline of a table on equally spaced knots as per-interval
        coefficients for the FP32 kernels: S = c0 + s (c1 + s (c2 + s c3)) with s = (x - args[i])
        / h in [0, 1] on interval i, derived in FP64 from the knots, values and second
        derivatives (Table's spline, src/Table.cpp:560-577) and rounded to float.  Returns the
        offset (floats) of the table's quadruples in the coefficient region of
        ``tables32_array``; the leaf records hold it, made absolute by ``leaves_array``."""
        if key not in self.coef_off:
            c = spline_coefficients(args, vals, y2)
            self.coef_off[key] = self.ncoef
            self.coefs.append(c.astype(np.float32).ravel())
            self.ncoef += c.size
        return self.coef_off[key]

    def _coef_base(self):
        """Offset (floats) of the coefficient region in tables32_array: after the float copy of
        the FP64 tables, rounded up to a 16-byte boundary."""
        return -(-max(self.ntab, 1) // 4) * 4

    def tables32_array(self):
        """The FP32 kernels' table buffer: the FP64 tables rounded to float, then the spline
        coefficient quadruples of add_spline_coefs."""
        t = np.concatenate(self.tables) if self.tables else np.zeros(1)
        base = self._coef_base()
        out = np.zeros(base + self.ncoef, dtype=np.float32)
        out[:len(t)] = t
        if self.coefs:
            out[base:] = np.concatenate(self.coefs)
        return out

    def exponential_cliprows(self, r0, gsparams, jac):
        """Host row ranges of the reference's Exponential clipping (FP64 path only)."""
        if self.precision != 'f64' or self.Nk is None:
            return -1
        out = np.zeros(2 * (self.Nk + 1), dtype=np.int32)
        j = None if jac is None else np.ascontiguousarray(np.asarray(jac, float).reshape(4))
        _lib.check(_lib.lib().gsmi_exponential_cliprows(
            float(r0), _lib.ptr(_lib.gsp_array(gsparams)), _lib.ptr(j), int(self.Nk),
            float(self.dk), _lib.ptr(out)), "Exponential clip rows")
        off = self.nclip
        self.clip.append(out)
        self.nclip += len(out)
        return off

    def kclip_rows(self, ksq_max, r0, jac):
        """Host row ranges of a radial leaf clipped at (k r0)^2 > ksq_max (FP64 path only)."""
        if self.precision != 'f64' or self.Nk is None:
            return -1
        out = np.zeros(2 * (self.Nk + 1), dtype=np.int32)
        j = None if jac is None else np.ascontiguousarray(np.asarray(jac, float).reshape(4))
        _lib.check(_lib.lib().gsmi_kclip_rows(float(ksq_max), float(r0), _lib.ptr(j),
                                              int(self.Nk), float(self.dk), _lib.ptr(out)),
                   "k clip rows")
        off = self.nclip
        self.clip.append(out)
        self.nclip += len(out)
        return off

    def ops_array(self):
        a = np.zeros(len(self.ops), dtype=_lib.OP_DTYPE)
        for i, (code, arg, x, y, z) in enumerate(self.ops):
            a[i] = (code, arg, x, y, z)
        return a

    def leaves_array(self):
        a = np.zeros(len(self.leaves), dtype=_lib.LEAF_DTYPE)
        if self.leaves:
            for name in LeafRecord._names:
                a[name] = [r.d[name] for r in self.leaves]
        if self.precision == 'f32':
            # Sersic leaves: coefficient offsets relative to the region -> absolute in tables32
            sel = (a['type'] == _lib.LEAF_SERSIC) & (a['cliprows'] >= 0)
            a['cliprows'][sel] += self._coef_base()
        return a


def check_stack(ops, begin=0, count=None, precision='f64'):
    """Depth of the post-order program against the kernels' register stack."""
    depth = best = 0
    end = len(ops) if count is None else begin + count
    for code, *_ in ops[begin:end]:
        if code == _lib.OP_LEAF:
            depth += 1
        elif code in (_lib.OP_MUL, _lib.OP_ADD):
            depth -= 1
        best = max(best, depth)
    limit = _lib.STACK_DEPTH[precision]
    if depth != 1 or best > limit:
        raise _lib.GalSimAMDError("profile tree too deep for the device program (depth %d > %d)"
                                  % (best, limit))


def hoist_axis_factors(ctx, op0):
    """Move the trailing separable factors of the program that starts at op0 -- PHASE and SCALE
    applied to the whole result, and a diagonal-Jacobian Box multiplied into it (the pixel) --
    behind the program, where the kernels evaluate them once per k column and row.
    Returns (op_count, axis_count)."""
    prog = ctx.ops[op0:]
    axis = []
    while True:
        if len(prog) > 1 and prog[-1][0] in (_lib.OP_PHASE, _lib.OP_SCALE):
            axis.append(prog.pop())
            continue
        if (len(prog) > 2 and prog[-1][0] == _lib.OP_MUL and prog[-2][0] == _lib.OP_LEAF):
            leaf = ctx.leaves[prog[-2][1]]
            if leaf['type'] == _lib.LEAF_BOX and leaf['separable']:
                axis.append(prog[-2])
                del prog[-2:]
                continue
        break
    ctx.ops[op0:] = prog + axis
    return len(prog), len(axis)


def plan_sizes(prof, bounds):
    """drawFFT_makeKImage size planning (galsim/gsobject.py:1927-1973) with image scale 1."""
    N = prof.getGoodImageSize(1.0)
    init = bounds._getinitargs()
    image_N = max(int(np.max(np.abs(init))) * 2, int(np.max(bounds.numpyShape())))
    N = max(N, image_N)
    N = _lib.lib().gsmi_good_fft_size(int(N))
    N = max(N, prof.gsparams.minimum_fft_size)
    dk = 2. * np.pi / (N * 1.0)
    maxk = prof.maxk
    if N * dk / 2 > maxk:
        Nk = N
    else:
        Nk = int(np.ceil(maxk / dk)) * 2
    if Nk > prof.gsparams.maximum_fft_size:
        warnings.warn("drawFFT requires a very large FFT (%d)" % Nk)
    return N, Nk, dk


class _Packer:
    """Packs numpy sections into one byte buffer with 256-byte aligned offsets."""

    def __init__(self):
        self.parts = []
        self.size = 0

    def add(self, arr):
        arr = np.ascontiguousarray(arr)
        off = self.size
        self.parts.append((off, arr.view(np.uint8).reshape(-1)))
        self.size += (arr.nbytes + _ALIGN - 1) // _ALIGN * _ALIGN
        return off

    def host(self):
        buf = np.zeros(max(self.size, _ALIGN), dtype=np.uint8)
        for off, b in self.parts:
            buf[off:off + len(b)] = b
        return buf


def _device_buffer(torch, host_bytes):
    t = torch.from_numpy(host_bytes)
    return t.to('cuda', non_blocking=False)


def draw_fft(prof, bounds, precision='f64', return_device=False):
    """Render ``prof`` (already in image coordinates) onto ``bounds`` (centred on the profile
    origin, as drawFFT expects).  Returns (pixels, added_flux)."""
    if precision not in ('f64', 'f32'):
        raise ValueError("precision must be 'f64' or 'f32'")
    lib = _lib.lib()
    N, Nk, dk = plan_sizes(prof, bounds)
    ctx = Lowering(precision, Nk, dk)
    prof._lower(ctx, None)
    nops, naxis = hoist_axis_factors(ctx, 0)
    check_stack(ctx.ops, 0, nops, precision)
    ny, nx = bounds.numpyShape()
    if not (-N // 2 <= bounds.xmin and bounds.xmax <= N // 2 - 1 and
            -N // 2 <= bounds.ymin and bounds.ymax <= N // 2 - 1):
        raise _lib.GalSimAMDError("output bounds %r do not fit the N=%d grid" % (bounds, N))
    rows = ny * (N // 2 + 1)
    stamp = np.zeros(1, dtype=_lib.STAMP_DTYPE)
    stamp[0] = (N, Nk, nx, ny, bounds.xmin, bounds.ymin, 0, nops, 0, naxis, 0, 0, rows, dk, 0,
                0)
    torch = _torch()
    pk = _Packer()
    o_stamp = pk.add(stamp)
    o_ops = pk.add(ctx.ops_array())
    o_leaves = pk.add(ctx.leaves_array())
    tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
    o_tab = pk.add(tables)
    o_tab32 = pk.add(ctx.tables32_array())
    clip = np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, dtype=np.int32)
    o_clip = pk.add(clip)
    dev = _device_buffer(torch, pk.host())
    base = dev.data_ptr()
    fdt = torch.float64 if precision == 'f64' else torch.float32
    scratch = torch.empty((rows + 3 * (Nk // 2) + 2) * 2, dtype=fdt, device='cuda')
    out = torch.empty((ny, nx), dtype=fdt, device='cuda')
    stream = torch.cuda.current_stream().cuda_stream
    if precision == 'f64':
        rc = lib.gsmi_draw_fft_f64(base + o_stamp, 1, N, ny, base + o_ops, base + o_leaves,
                                   base + o_tab, base + o_clip, scratch.data_ptr(),
                                   out.data_ptr(), stream)
        if rc == -5:
            # an FFT size beyond the fused kernels' LDS: the same steps unfused on the GPU
            return _draw_fft_unfused(prof, bounds, N, Nk, dk, return_device)
    else:
        rc = lib.gsmi_draw_fft_f32(base + o_stamp, 1, N, ny, base + o_ops, base + o_leaves,
                                   base + o_tab, base + o_tab32, scratch.data_ptr(),
                                   out.data_ptr(), stream)
    _lib.check(rc, "gsmi_draw_fft_%s" % precision)
    if return_device:
        return out, (dev, scratch)
    host = out.cpu().numpy()
    return host, float(host.sum(dtype=float))


def _draw_fft_unfused(prof, bounds, N, Nk, dk, return_device=False):
    """drawFFT in FP64 as the reference's separate steps, for FFT sizes whose column transform
    does not fit one workgroup's LDS (N = 6144, 8192): the k-image on (0..Nk/2, -Nk/2..Nk/2)
    (gsmi_draw_kimage_f64, with the reference's clip rows), wrapImage(hermx) onto the N-grid
    (gsmi_wrap_hermx_f64), irfft(shift_in, shift_out) (gsmi_irfft2_f64, whose column transform
    is then four-step) and the crop (galsim/gsobject.py:1927-2045)."""
    from .fft import irfft_scratch_rows
    lib = _lib.lib()
    ctx = Lowering('f64', Nk, dk)
    prof._lower(ctx, None)
    check_stack(ctx.ops, precision='f64')
    torch = _torch()
    pk = _Packer()
    o_ops = pk.add(ctx.ops_array())
    o_leaves = pk.add(ctx.leaves_array())
    tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
    o_tab = pk.add(tables)
    clip = np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, dtype=np.int32)
    o_clip = pk.add(clip)
    dev = _device_buffer(torch, pk.host())
    base = dev.data_ptr()
    stream = torch.cuda.current_stream().cuda_stream
    h = Nk // 2
    kimg = torch.empty((Nk + 1, h + 1), dtype=torch.complex128, device='cuda')
    _lib.check(lib.gsmi_draw_kimage_f64(base + o_ops, len(ctx.ops), base + o_leaves, base + o_tab,
                                        base + o_clip, 0, -h, h + 1, Nk + 1, Nk, float(dk),
                                        kimg.data_ptr(), stream), "gsmi_draw_kimage_f64")
    kw = torch.empty((N, N // 2 + 1), dtype=torch.complex128, device='cuda')
    _lib.check(lib.gsmi_wrap_hermx_f64(kimg.data_ptr(), Nk, N, kw.data_ptr(), stream),
               "gsmi_wrap_hermx_f64")
    del kimg
    real = torch.empty((N, N), dtype=torch.float64, device='cuda')
    tmp = torch.empty((irfft_scratch_rows(N), N // 2 + 1), dtype=torch.complex128, device='cuda')
    _lib.check(lib.gsmi_irfft2_f64(kw.data_ptr(), N, N, 1, 1, real.data_ptr(), tmp.data_ptr(),
                                   stream), "gsmi_irfft2_f64")
    del kw, tmp
    out = real[bounds.ymin + N // 2: bounds.ymax + N // 2 + 1,
               bounds.xmin + N // 2: bounds.xmax + N // 2 + 1].contiguous()
    if return_device:
        return out, (dev,)
    host = out.cpu().numpy()
    return host, float(host.sum(dtype=float))


def draw_kimage(prof, bounds, dk, jac=None, Nk=None, return_device=False):
    """k-image of ``prof`` on ``bounds`` with spacing dk (complex128): SBProfile::drawK
    (src/SBProfile.cpp:383-425), k mapped by the 2x2 Jacobian ``jac`` if given.  With ``Nk`` the
    bounds are drawFFT_makeKImage's (0..Nk/2, -Nk/2..Nk/2) and Exponential / Spergel leaves clip
    rows exactly as the reference's GetKValueRange1d/2d does on that image; otherwise they clip
    at k^2 > ksq_max.  Returns a host array, or with return_device a (ny, nx, 2) float64 device
    tensor."""
    lib = _lib.lib()
    if Nk is not None and not (bounds.xmin == 0 and bounds.xmax == Nk // 2 and
                               bounds.ymin == -(Nk // 2) and bounds.ymax == Nk // 2):
        raise _lib.GalSimAMDError("clip rows need the bounds (0..Nk/2, -Nk/2..Nk/2)")
    ctx = Lowering('f64', Nk=Nk, dk=dk)
    ny, nx = bounds.numpyShape()
    ctx.kbounds = (bounds.xmin, bounds.ymin, nx, ny)
    prof._lower(ctx, None if jac is None else np.asarray(jac, dtype=float).reshape(2, 2))
    check_stack(ctx.ops, precision='f64')
    ny, nx = bounds.numpyShape()
    torch = _torch()
    pk = _Packer()
    o_ops = pk.add(ctx.ops_array())
    o_leaves = pk.add(ctx.leaves_array())
    tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
    o_tab = pk.add(tables)
    clip = np.concatenate(ctx.clip) if ctx.clip else np.zeros(2, dtype=np.int32)
    o_clip = pk.add(clip)
    dev = _device_buffer(torch, pk.host())
    base = dev.data_ptr()
    out = torch.empty((ny, nx, 2), dtype=torch.float64, device='cuda')
    rc = lib.gsmi_draw_kimage_f64(base + o_ops, len(ctx.ops), base + o_leaves, base + o_tab,
                                  base + o_clip, bounds.xmin, bounds.ymin, nx, ny,
                                  int(Nk or 0), float(dk), out.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, "gsmi_draw_kimage_f64")
    if return_device:
        return out
    h = out.cpu().numpy()
    return h[..., 0] + 1j * h[..., 1]


def kvalue(prof, kx, ky):
    """k-space value at (kx, ky): the program with the diagonal Jacobian diag(kx, ky) at the
    grid point (1, 1) with dk = 1."""
    lib = _lib.lib()
    ctx = Lowering('f64')
    ctx.kvalue_mode = True
    prof._lower(ctx, np.array([[float(kx), 0.], [0., float(ky)]]))
    check_stack(ctx.ops, precision='f64')
    torch = _torch()
    pk = _Packer()
    o_ops = pk.add(ctx.ops_array())
    o_leaves = pk.add(ctx.leaves_array())
    tables = np.concatenate(ctx.tables) if ctx.tables else np.zeros(1)
    o_tab = pk.add(tables)
    dev = _device_buffer(torch, pk.host())
    base = dev.data_ptr()
    out = torch.empty(2, dtype=torch.float64, device='cuda')
    rc = lib.gsmi_draw_kimage_f64(base + o_ops, len(ctx.ops), base + o_leaves, base + o_tab,
                                  None, 1, 1, 1, 1, 0, 1.0, out.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, "gsmi_draw_kimage_f64")
    h = out.cpu().numpy()
    return complex(h[0], h[1])


def irfft(kwrap, N):
    """irfft(shift_in, shift_out) of a wrapped (N, N/2+1) complex k-image (rows ky=-N/2..N/2-1)
    -> (N, N) real image on (-N/2..N/2-1)^2, FP64, on the GPU."""
    lib = _lib.lib()
    torch = _torch()
    k = torch.from_numpy(np.ascontiguousarray(kwrap, dtype=np.complex128).view(np.float64)).cuda()
    scratch = torch.empty(N * (N // 2 + 1) * 2, dtype=torch.float64, device='cuda')
    out = torch.empty((N, N), dtype=torch.float64, device='cuda')
    _lib.check(lib.gsmi_irfft_f64(k.data_ptr(), int(N), out.data_ptr(), scratch.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream), "gsmi_irfft_f64")
    return out.cpu().numpy()


def wrap_hermx(karr, Nk, N):
    """wrapImage(hermx=True) of a (Nk+1, Nk/2+1) k-image onto (N, N/2+1), on the GPU."""
    lib = _lib.lib()
    torch = _torch()
    k = torch.from_numpy(np.ascontiguousarray(karr, dtype=np.complex128).view(np.float64)).cuda()
    out = torch.empty((N, N // 2 + 1, 2), dtype=torch.float64, device='cuda')
    _lib.check(lib.gsmi_wrap_hermx_f64(k.data_ptr(), int(Nk), int(N), out.data_ptr(),
                                       torch.cuda.current_stream().cuda_stream),
               "gsmi_wrap_hermx_f64")
    h = out.cpu().numpy()
    return h[..., 0] + 1j * h[..., 1]


def _check_fft_image(image, what):
    if image.scale is None:
        raise _lib.GalSimAMDError("%s requires that the scale be set." % what)
    if not image.bounds.isDefined():
        raise _lib.GalSimAMDUndefinedBoundsError(
            "%s requires that the image have defined bounds." % what)


def forward_fft(image):
    """Image.calculate_fft (galsim/image.py:915-956): the image zero-padded to
    (-N/2..N/2-1)^2, N/2 = max(-xmin, xmax+1, -ymin, ymax+1), transformed by
    rfft(shift_in=True, shift_out=True) on the GPU and scaled by dx^2; the result lies on
    (0..N/2, -N/2..N/2-1) with scale 2 pi / (N dx).  A complex image contributes its real part, as
    the reference's rfft reads it."""
    from .image import Image
    _check_fft_image(image, "calculate_fft")
    b = image.bounds
    No2 = max(-b.xmin, b.xmax + 1, -b.ymin, b.ymax + 1)
    N = 2 * No2
    arr = np.asarray(image.array)
    if arr.dtype.kind == 'c':
        arr = arr.real
    full = np.zeros((N, N), dtype=np.float64)
    full[b.ymin + No2: b.ymax + No2 + 1, b.xmin + No2: b.xmax + No2 + 1] = arr
    torch = _torch()
    x = torch.from_numpy(full).cuda()
    out = torch.empty((N, No2 + 1), dtype=torch.complex128, device='cuda')
    tmp = torch.empty((N, No2 + 1), dtype=torch.complex128, device='cuda')
    _lib.check(_lib.lib().gsmi_rfft2_f64(x.data_ptr(), N, N, 1, 1, out.data_ptr(),
                                         tmp.data_ptr(), torch.cuda.current_stream().cuda_stream),
               "gsmi_rfft2_f64")
    dx = image.scale
    karr = out.cpu().numpy()
    karr *= dx * dx
    kim = Image(karr, xmin=0, ymin=-No2)
    kim.scale = np.pi / (No2 * dx)
    return kim


def inverse_fft(kimage):
    """Image.calculate_inverse_fft (galsim/image.py:958-1016): the x >= 0 half of the k-image,
    taken as Hermitian, zero-padded and wrapped (hermitian='x') onto (0..N/2, -N/2..N/2-1) unless
    it already has exactly those bounds, N/2 = max(xmax, -ymin, ymax); then
    irfft(shift_in=True, shift_out=True) on the GPU, scaled by (dk N / 2 pi)^2, on
    (-N/2..N/2-1)^2 with scale 2 pi / (N dk)."""
    from .image import Image
    from .position import BoundsI
    _check_fft_image(kimage, "calculate_inverse_fft")
    b = kimage.bounds
    if not b.includes(0, 0):
        raise _lib.GalSimAMDBoundsError(
            "calculate_inverse_fft requires that the image includes (0,0)")
    No2 = max(b.xmax, -b.ymin, b.ymax)
    if No2 < 1:
        raise _lib.GalSimAMDBoundsError("calculate_inverse_fft needs more than the (0,0) pixel")
    N = 2 * No2
    torch = _torch()
    lib = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    arr = np.asarray(kimage.array).astype(np.complex128)
    if b == BoundsI(0, No2, -No2, No2 - 1):
        kw = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
    else:
        full = np.zeros((N + 1, No2 + 1), dtype=np.complex128)
        full[b.ymin + No2: b.ymax + No2 + 1, 0: b.xmax + 1] = arr[:, -b.xmin:]
        kf = torch.from_numpy(full).cuda()
        kw = torch.empty((N, No2 + 1), dtype=torch.complex128, device='cuda')
        _lib.check(lib.gsmi_wrap_hermx_f64(kf.data_ptr(), N, N, kw.data_ptr(), stream),
                   "gsmi_wrap_hermx_f64")
    from .fft import irfft_scratch_rows
    out = torch.empty((N, N), dtype=torch.float64, device='cuda')
    tmp = torch.empty((irfft_scratch_rows(N), No2 + 1), dtype=torch.complex128, device='cuda')
    _lib.check(lib.gsmi_irfft2_f64(kw.data_ptr(), N, N, 1, 1, out.data_ptr(), tmp.data_ptr(),
                                   stream), "gsmi_irfft2_f64")
    dk = kimage.scale
    real = out.cpu().numpy()
    real *= (dk * No2 / np.pi) ** 2
    im = Image(real, xmin=-No2, ymin=-No2)
    im.scale = np.pi / (No2 * dk)
    return im
