"""Oracle numerics: Brent root finder, natural cubic spline table, goodFFTSize, Ogata Hankel.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import math

import numpy as np
import scipy.special as sps

DBL_EPS = np.finfo(float).eps


# ----------------------------------------------------------------------------------------------
# fmath::expd -- include/galsim/fmath/fmath.hpp:157-192, 457-497
# ----------------------------------------------------------------------------------------------
_EXPD_SBIT = 11
_EXPD_S = 1 << _EXPD_SBIT
_EXPD_A = _EXPD_S / math.log(2.0)
_EXPD_RA = 1. / _EXPD_A
_EXPD_ADJ = (1 << (_EXPD_SBIT + 10)) - (1 << _EXPD_SBIT)
_EXPD_C1 = 1.0
_EXPD_C2 = 0.16666666685227835064
_EXPD_C3 = 3.0000000027955394
_EXPD_TBL = np.array([np.float64(math.pow(2.0, i * (1.0 / _EXPD_S))).view(np.uint64)
                      & np.uint64((1 << 52) - 1) for i in range(_EXPD_S)], dtype=np.uint64)
_EXPD_B = float(3 << 51)


def expd(x):
    """The table-driven exp of the fmath library the reference vendors (fmath.hpp:457-497):
    2^(n/2048) from a 2048-entry mantissa table times a cubic in the remainder.  Used wherever the
    reference calls ``fmath::expd``; it differs from a correctly rounded exp by ~1 ulp, which
    matters only where the reference amplifies it (the Sersic high-k fit).  Pinned bit for bit
    against oracle/_ref (``SersicInfo::kValue`` in the high-k region)."""
    x = np.asarray(x, dtype=np.float64)
    scalar = x.ndim == 0
    x = np.atleast_1d(x)
    with np.errstate(invalid='ignore', over='ignore'):
        xc = np.clip(np.nan_to_num(x, nan=0.), -708.4, 709.8)
        d = xc * _EXPD_A + _EXPD_B
        di = (d.view(np.uint64) & np.uint64(0xffffffff)).astype(np.int64)
        di = np.where(di >= (1 << 31), di - (1 << 32), di)          # int32 -> sign-extended
        iax = _EXPD_TBL[(di & (_EXPD_S - 1)).astype(np.int64)]
        t = (d - _EXPD_B) * _EXPD_RA - xc
        u = (((di + _EXPD_ADJ) >> _EXPD_SBIT) << 52).astype(np.uint64)
        y = (_EXPD_C3 - t) * (t * t) * _EXPD_C2 - t + _EXPD_C1
        did = (u | iax).view(np.float64)
        out = y * did
    out = np.where(np.isnan(x), np.nan, out)
    out = np.where(x <= -708.39641853226408, 0., out)
    out = np.where(x >= 709.78271289338397, np.inf, out)
    return float(out[0]) if scalar else out


# ----------------------------------------------------------------------------------------------
# goodFFTSize -- src/Image.cpp:1139-1152
# ----------------------------------------------------------------------------------------------
def good_fft_size(n):
    """Smallest 2^k or 3*2^k (>= 6) that is >= n, with the reference's 1e-5 slops.

    Restates ``goodFFTSize`` (src/Image.cpp:1141-1152).  Pinned by tests/test_draw.py:1306
    (``good_fft_size(1174) == 1536``) and by brute force in tests/test_oracle_numerics.py.
    """
    n = int(n)
    if n <= 2:
        return 2
    insize = (1. - 1.e-5) * n
    log2n = math.log(2.) * math.ceil(math.log(insize) / math.log(2.))
    log2n3 = math.log(3.) + math.log(2.) * math.ceil((math.log(insize) - math.log(3.)) /
                                                      math.log(2.))
    log2n3 = max(log2n3, math.log(6.))
    return int(math.ceil(math.exp(min(log2n, log2n3)) - 1.e-5))


# ----------------------------------------------------------------------------------------------
# Brent solver -- include/galsim/Solve.h
# ----------------------------------------------------------------------------------------------
class SolveError(RuntimeError):
    pass


class Solve:
    """Restatement of ``galsim::Solve`` (include/galsim/Solve.h:124-466) with the Brent method.

    Only the members on the FFT path are restated: ``bracketUpper`` (bracket1, Solve.h:213-248),
    ``bracketLowerWithLimit`` (bracket1WithLimit, Solve.h:264-308) and ``zbrent``
    (Solve.h:396-462).  Default x tolerance 1e-7 and 40 steps (Solve.h:45-46).
    Pinned against scipy.optimize.brentq roots (tests/test_oracle_numerics.py) and, through the
    Sersic/Moffat scalars, against oracle/_ref.
    """

    def __init__(self, func, lb, ub, xtol=1.e-7, max_steps=40):
        self.func = func
        self.lb = float(lb)
        self.ub = float(ub)
        self.xtol = xtol
        self.max_steps = max_steps
        self._evaluated = False

    def _evaluate_bounds(self):
        if not self._evaluated:
            self.flower = self.func(self.lb)
            self.fupper = self.func(self.ub)
            self._evaluated = True

    def bracket_upper(self):
        # Solve.h:213-248 (bracket1 with a = lBound, b = uBound; b aliases uBound).
        if self.ub == self.lb:
            raise SolveError("uBound=lBound in bracketUpper()")
        self._evaluate_bounds()
        delta = self.ub - self.lb
        for _ in range(1, self.max_steps):
            if self.fupper * self.flower <= 0.:
                return
            self.lb = self.ub
            self.flower = self.fupper
            delta *= 2.
            self.ub += delta
            self.fupper = self.func(self.ub)
        raise SolveError("Too many iterations in bracketUpper()")

    def bracket_lower_with_limit(self, lower_limit):
        # Solve.h:264-308 (bracket1WithLimit with a = uBound, b = lBound, c = lower_limit).
        if self.ub == self.lb:
            raise SolveError("uBound=lBound in bracketLowerWithLimit()")
        if self.lb == lower_limit:
            raise SolveError("lBound=lower_limit in bracketLowerWithLimit()")
        if (self.ub - self.lb) * (self.lb - lower_limit) <= 0:
            raise SolveError("lBound not between uBound and lower_limit")
        self._evaluate_bounds()
        c = lower_limit
        for _ in range(1, self.max_steps):
            if self.fupper * self.flower <= 0.:
                return
            bma = self.lb - self.ub
            cmb = c - self.lb
            self.ub = self.lb
            self.fupper = self.flower
            self.lb = (2. * c / cmb + self.ub / bma) / (2. / cmb + 1. / bma)
            self.flower = self.func(self.lb)
        raise SolveError("Too many iterations in bracketLowerWithLimit()")

    def bisect(self):
        # bisect, Solve.h:376-393 (the method a Solve uses unless setMethod(Brent) is called)
        self._evaluate_bounds()
        f, fmid = self.flower, self.fupper
        if f * fmid > 0.:
            raise SolveError("Root is not bracketed")
        if f < 0.:
            dx, rtb = self.ub - self.lb, self.lb
        else:
            dx, rtb = self.lb - self.ub, self.ub
        for _ in range(self.max_steps):
            dx *= 0.5
            xmid = rtb + dx
            fmid = self.func(xmid)
            if fmid <= 0.:
                rtb = xmid
            if abs(dx) < self.xtol or fmid == 0.:
                return rtb
        raise SolveError("Too many bisections")

    def root(self):
        # zbrent, Solve.h:396-462
        a = self.lb
        b = self.ub
        c = self.ub
        d = b - a
        e = b - a
        self._evaluate_bounds()
        fa = self.flower
        fb = self.fupper
        if (fa > 0. and fb > 0.) or (fa < 0. and fb < 0.):
            raise SolveError("Root is not bracketed: %r %r" % (self.lb, self.ub))
        fc = fb
        for _ in range(self.max_steps + 1):
            if (fb > 0. and fc > 0.) or (fb < 0. and fc < 0.):
                c = a
                fc = fa
                e = d = b - a
            if abs(fc) < abs(fb):
                a = b
                b = c
                c = a
                fa = fb
                fb = fc
                fc = fa
            tol1 = 2.0 * DBL_EPS * abs(b) + 0.5 * self.xtol
            xm = 0.5 * (c - b)
            if abs(xm) <= tol1 or fb == 0.0:
                return b
            if abs(e) >= tol1 and abs(fa) > abs(fb):
                s = fb / fa
                if a == c:
                    p = 2.0 * xm * s
                    q = 1.0 - s
                else:
                    q = fa / fc
                    r = fb / fc
                    p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0))
                    q = (q - 1.0) * (r - 1.0) * (s - 1.0)
                if p > 0.0:
                    q = -q
                p = abs(p)
                min1 = 3.0 * xm * q - abs(tol1 * q)
                min2 = abs(e * q)
                if 2.0 * p < min(min1, min2):
                    e = d
                    d = p / q
                else:
                    d = xm
                    e = d
            else:
                d = xm
                e = d
            a = b
            fa = fb
            if abs(d) > tol1:
                b += d
            else:
                b += abs(tol1) if xm >= 0. else -abs(tol1)
            fb = self.func(b)
        raise SolveError("Maximum number of iterations exceeded in zbrent")


# ----------------------------------------------------------------------------------------------
# Natural cubic spline table -- src/Table.cpp
# ----------------------------------------------------------------------------------------------
class SplineTable:
    """Natural cubic spline lookup table (``Table::spline``).

    Second derivatives by the Thomas algorithm exactly as ``TSpline::setupSpline``
    (src/Table.cpp:635-686); index search as ``ArgVec::upperIndex`` (src/Table.cpp:62-120,
    equal-spaced fast path with the 0.01 tolerance test of 64-74); interpolation as
    ``TSpline::_interp`` (src/Table.cpp:560-577); out-of-range -> 0 as ``Table::operator()``
    (src/Table.cpp:780-785).  Pinned against scipy.interpolate.CubicSpline(bc_type='natural')
    in tests/test_oracle_numerics.py.
    """

    def __init__(self, args, vals):
        self.args = np.asarray(args, dtype=float)
        self.vals = np.asarray(vals, dtype=float)
        n = len(self.args)
        self.n = n
        y2 = np.zeros(n)
        a_, v_ = self.args, self.vals
        if n == 3:
            y2[1] = 3. * ((v_[2] - v_[1]) / (a_[2] - a_[1]) -
                          (v_[1] - v_[0]) / (a_[1] - a_[0])) / (a_[2] - a_[0])
        elif n > 3:
            c = np.zeros(n - 3)
            for i in range(1, n - 1):
                y2[i] = 6. * ((v_[i + 1] - v_[i]) / (a_[i + 1] - a_[i]) -
                              (v_[i] - v_[i - 1]) / (a_[i] - a_[i - 1]))
            bb = 2. * (a_[2] - a_[0])
            for i in range(1, n - 1):
                y2[i] /= bb
                if i == n - 2:
                    break
                a = a_[i + 1] - a_[i]
                c[i - 1] = a
                c[i - 1] /= bb
                bb = 2. * (a_[i + 2] - a_[i])
                bb -= a * c[i - 1]
                y2[i + 1] -= a * y2[i]
            for i in range(n - 3, 0, -1):
                y2[i] -= c[i - 1] * y2[i + 1]
        self.y2 = y2
        # ArgVec constructor, src/Table.cpp:62-74
        self.da = (a_[-1] - a_[0]) / (n - 1)
        self.equal_spaced = bool(np.all(np.abs((a_[1:] - a_[0]) / self.da -
                                               np.arange(1, n)) <= 0.01))

    @property
    def arg_min(self):
        return self.args[0]

    @property
    def arg_max(self):
        return self.args[-1]

    def upper_index(self, a):
        """Vectorised ArgVec::upperIndex (src/Table.cpp:77-120)."""
        a = np.asarray(a, dtype=float)
        n = self.n
        if self.equal_spaced:
            i = np.ceil((a - self.args[0]) / self.da).astype(np.int64)
            i = np.clip(i, 1, n - 1)
        else:
            i = np.searchsorted(self.args, a, side='left')
            i = np.clip(i, 1, n - 1)
        i = np.where(a < self.args[0], 1, i)
        i = np.where(a > self.args[-1], n - 1, i)
        return i

    def lookup(self, a):
        a = np.asarray(a, dtype=float)
        i = self.upper_index(a)
        h = self.args[i] - self.args[i - 1]
        aa = self.args[i] - a
        bb = h - aa
        return (aa * self.vals[i - 1] + bb * self.vals[i] -
                (1. / 6.) * aa * bb * ((aa + h) * self.y2[i - 1] + (bb + h) * self.y2[i])) / h

    def __call__(self, a):
        a = np.asarray(a, dtype=float)
        out = self.lookup(a)
        return np.where((a < self.arg_min) | (a > self.arg_max), 0., out)


# ----------------------------------------------------------------------------------------------
# Ogata (2005) Hankel transform -- src/math/Hankel.cpp:79-247
# ----------------------------------------------------------------------------------------------
_J0_ZEROS = None


def _j0_zeros(n):
    global _J0_ZEROS
    if _J0_ZEROS is None or len(_J0_ZEROS) < n:
        _J0_ZEROS = sps.jn_zeros(0, max(n, 4096))
    return _J0_ZEROS[:n]


class _HankelIntegrator:
    """One Ogata quadrature rule for order 0 and step h (src/Hankel.cpp:79-181).

    Nodes ``x_i = pi/h psi(h xi_i)``, ``xi_i = j_{0,i}/pi``, weights
    ``Y0(pi xi)/J1(pi xi) * pi x_i J0(x_i) psi'(h xi)``; summation stops at the first term with
    ``|step| < 1e-15 |ans|``, weights are produced in batches of 256 up to ``min(pi/h, 1e6)``.
    """

    def __init__(self, h, batch=256):
        self.h = h
        self.nmax = int(min(math.pi / h, 1.e6))
        self.batch = batch
        self.x = np.zeros(0)
        self.w = np.zeros(0)
        self._extend()

    @staticmethod
    def _psi(t):
        return t * np.tanh(np.pi / 2. * np.sinh(t))

    def _dpsi(self, t):
        return (t * np.pi / 2. * np.cosh(t) / np.cosh(np.pi / 2. * np.sinh(t)) ** 2 +
                self._psi(t) / t)

    def _extend(self):
        n1 = len(self.x)
        n2 = min(n1 + self.batch, self.nmax)
        xi = _j0_zeros(n2)[n1:n2] / np.pi
        t = self.h * xi
        x = np.pi / self.h * self._psi(t)
        w = sps.y0(np.pi * xi) / sps.j1(np.pi * xi)
        w = w * (np.pi * x * sps.j0(x) * self._dpsi(t))
        self.x = np.concatenate([self.x, x])
        self.w = np.concatenate([self.w, w])

    def integrate(self, f, k):
        ans = 0.
        n1 = 0
        while True:
            steps = self.w[n1:] * f(self.x[n1:] / k)
            # np.cumsum accumulates sequentially, i.e. the same rounding as "ans += step".
            run = np.cumsum(np.concatenate([[ans], steps]))[1:]
            stop = np.nonzero(np.abs(steps) < 1.e-15 * np.abs(run))[0]
            if len(stop):
                return run[stop[0]] / (k * k)
            ans = run[-1]
            n1 = len(self.x)
            if len(self.x) == self.nmax or steps[-1] == 0:
                return ans / (k * k)
            self._extend()


class _AdaptiveHankel:
    """Halve h from 1/32 until two successive rules agree (src/math/Hankel.cpp:183-229)."""

    def __init__(self, h0=1. / 32.):
        self.h0 = h0
        self.rules = {}

    def _rule(self, h):
        if h not in self.rules:
            self.rules[h] = _HankelIntegrator(h)
        return self.rules[h]

    def integrate(self, f, k, relerr, abserr):
        h0 = self.h0
        while h0 > 100 * k:
            h0 *= 0.5
        ans0 = self._rule(h0).integrate(f, k)
        h1 = 0.5 * h0
        ans1 = self._rule(h1).integrate(f, k)
        err = abs(ans1 - ans0)
        while ((err > relerr * ans1 and (err > abserr or abs(ans1) > 2 * abs(ans0)))
               or ans1 == 0.):
            h0 = h1
            ans0 = ans1
            h1 *= 0.5
            ans1 = self._rule(h1).integrate(f, k)
            err = abs(ans1 - ans0)
        return ans1


_HANKEL0 = _AdaptiveHankel()


def hankel_inf(f, k, relerr, abserr):
    """Order-0 Hankel transform int_0^inf f(r) J0(k r) r dr (src/math/Hankel.cpp:231-247).

    ``f`` must accept numpy arrays.  k == 0 (GKP quadrature in the reference) is not needed on
    the FFT path and is not restated.  Pinned against oracle/_ref ``hankel_inf`` and against the
    closed form of the exponential's transform in tests/test_oracle_numerics.py.
    """
    if k == 0.:
        raise NotImplementedError("hankel_inf at k=0 is not on the FFT path")
    return _HANKEL0.integrate(f, k, relerr, abserr)
