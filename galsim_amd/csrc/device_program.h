// Device-side program evaluation shared by the forward and adjoint kernels: the leaf k-space
// formulas, the register-stack program of include/gsmi.h, the hoisted axis factors, and the
// in-LDS Stockham FFT.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/gsmi.h"
#include "device_math.h"

namespace gsmi {

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------------------------
// Leaf evaluation
// ---------------------------------------------------------------------------------------------
template <typename T>
struct Tables;

template <>
struct Tables<double> {
    const double* tab;
    const int32_t* clip;
};
template <>
struct Tables<float> {
    const double* tab;
    const float* tab32;
    const float4* coef;   // Sersic cubic coefficients (LeafI::cf indexes it); see sersic_f32
};

// TopHat's 2 J1(x) / x (x^2 >= 1e-4), out of line: the Bessel function's code would otherwise
// be inlined into every fold kernel's interpreter path and raise its register count.
__device__ __noinline__ double tophat_j1_ratio(double x) { return 2. * j1(x) / x; }
__device__ __noinline__ float tophat_j1_ratio(float x) { return 2.f * j1f(x) / x; }

// Moffat kV for the Bessel-function cases, kept out of line so the common path stays small.
__device__ __noinline__ double moffat_kv_bessel(int kind, double nu, double k, double ksq,
                                                double at_zero)
{
    switch (kind) {
      case 2: return ksq == 0. ? 1. : bessel_k(1., k) * k;
      case 3: return ksq == 0. ? 2. : bessel_k(2., k) * ksq;
      case 4: return ksq == 0. ? 8. : bessel_k(3., k) * k * ksq;
      default: return ksq == 0. ? at_zero : bessel_k(nu, k) * pow(k, nu);
    }
}

// A leaf decoded once per stamp: its constants converted to the arithmetic type (LeafC, kept in
// LDS by the batched kernels) and its integer fields (LeafI, uniform registers).
template <typename T>
struct LeafC {
    T j0, j1, j2, j3;     // composite Jacobian (dudx, dudy, dvdx, dvdy)
    T s, s2, flux, t0;    // scale, Box height scale, flux (Moffat: knorm); t0..t2: Sersic
    T p[6];               // type-specific constants
    T t1, t2;             // Sersic: t0, t1 = first and last log-k knot, t2 = 1 / knot spacing
    T q0, q1, q2, q3;     // |k'|^2 s^2 = q0 kx^2 + q1 kx ky + q2 ky^2 (q3: padding)
};

struct LeafI {
    int type, kind, sep, aux, ntab;   // aux: Sersic table offset, or Exponential clip-row offset
    int cf;                           // Sersic (FP32): index of its first cubic coefficient
                                      // quadruple in Tables<float>::coef, -1 if none
};

template <typename T>
__device__ inline LeafC<T> leaf_consts(const gsmi_leaf& L)
{
    LeafC<T> c;
    c.j0 = T(L.jac[0]);
    c.j1 = T(L.jac[1]);
    c.j2 = T(L.jac[2]);
    c.j3 = T(L.jac[3]);
    c.s = T(L.scale);
    c.s2 = T(L.scale2);
    c.flux = T(L.flux);
    c.t0 = c.t1 = c.t2 = T(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) c.p[k] = T(L.p[k]);
    // the quadratic form of k' = J^T k scaled by s, in FP64 from the record
    const double s2 = L.scale * L.scale;
    c.q0 = T(s2 * (L.jac[0] * L.jac[0] + L.jac[1] * L.jac[1]));
    c.q1 = T(2. * s2 * (L.jac[0] * L.jac[2] + L.jac[1] * L.jac[3]));
    c.q2 = T(s2 * (L.jac[2] * L.jac[2] + L.jac[3] * L.jac[3]));
    c.q3 = T(0);
    return c;
}

__device__ inline LeafI leaf_ints(const gsmi_leaf& L)
{
    LeafI r;
    r.type = L.type;
    r.kind = L.kind;
    r.sep = L.separable;
    r.aux = (L.type == GSMI_LEAF_SERSIC || (L.type == GSMI_LEAF_MOFFAT && L.kind == 1))
                ? L.table : L.cliprows;
    r.ntab = L.ntab;
    r.cf = (L.type == GSMI_LEAF_SERSIC && L.cliprows >= 0) ? L.cliprows / 4 : -1;
    return r;
}

// Table::lookup of the natural cubic spline on equally spaced log-k knots (src/Table.cpp:77-92,
// 560-577), with its derivative when dS != nullptr.
template <typename T>
__device__ inline T spline_eq(const T* __restrict__ args, int n, T lk, T* dS)
{
    const T* vals = args + n;
    const T* y2 = vals + n;
    const T a0 = args[0];
    if (lk < a0 || lk > args[n - 1]) {
        if (dS) *dS = T(0);
        return T(0);
    }
    int i;
    if constexpr (sizeof(T) == 8) {
        const double da = (args[n - 1] - a0) / (n - 1);
        i = int(ceil((lk - a0) / da));
    } else {
        const float inv_da = float(n - 1) / (args[n - 1] - a0);
        i = int(ceilf((lk - a0) * inv_da));
    }
    i = min(max(i, 1), n - 1);
    const T h = args[i] - args[i - 1];
    const T aa = args[i] - lk;
    const T bb = h - aa;
    const T cc = (aa + h) * y2[i - 1] + (bb + h) * y2[i];
    if (dS)
        *dS = ddiv(vals[i] - vals[i - 1] -
                   (T(1) / T(6)) * ((aa - bb) * cc + aa * bb * (y2[i] - y2[i - 1])), h);
    return ddiv(aa * vals[i - 1] + bb * vals[i] - (T(1) / T(6)) * aa * bb * cc, h);
}

template <typename T>
__device__ inline const T* sersic_table(const Tables<T>& tb, int off);
template <>
__device__ inline const double* sersic_table<double>(const Tables<double>& tb, int off)
{
    return tb.tab + off;
}
template <>
__device__ inline const float* sersic_table<float>(const Tables<float>& tb, int off)
{
    return tb.tab32 + off;
}

// A wave-uniform value moved to a scalar register (the compiler cannot prove LDS loads uniform).
__device__ __forceinline__ float uniform(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double uniform(double x)
{
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_readfirstlane(int(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane(int(b >> 32));
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) |
                                          static_cast<long long>(static_cast<unsigned>(lo)));
}

// A decoded leaf copied out of LDS into scalar registers: the sample loops then read its
// constants as instruction operands instead of reloading them from LDS.
template <typename T>
__device__ __forceinline__ LeafC<T> uniform_leaf(const LeafC<T>& in)
{
    LeafC<T> out;
    out.j0 = uniform(in.j0); out.j1 = uniform(in.j1); out.j2 = uniform(in.j2);
    out.j3 = uniform(in.j3); out.s = uniform(in.s); out.s2 = uniform(in.s2);
    out.flux = uniform(in.flux); out.t0 = uniform(in.t0);
#pragma unroll
    for (int k = 0; k < 6; ++k) out.p[k] = uniform(in.p[k]);
    out.t1 = uniform(in.t1); out.t2 = uniform(in.t2);
    out.q0 = uniform(in.q0); out.q1 = uniform(in.q1); out.q2 = uniform(in.q2);
    out.q3 = T(0);
    return out;
}

// Truncated Moffat kV_trunc (src/SBMoffat.cpp:340-345): the natural spline of its Hankel table
// on x = (k rD)^2, knots at the squares of k = 0, dk, 2 dk, ... (a running sum), zero beyond the
// last knot.  The interval comes from k / dk, stepped to the one holding x (the reference's
// search gives it, or its neighbour when x is a knot, where the spline is continuous).  With dS,
// also dS/dx.
template <typename T>
__device__ inline T moffat_trunc_spline(const T* __restrict__ args, int n, T x, T k, T inv_dk,
                                        T* dS)
{
    if (x > args[n - 1]) {
        if (dS) *dS = T(0);
        return T(0);
    }
    int i = int(ceil(k * inv_dk));
    i = min(max(i, 1), n - 1);
    if (i < n - 1 && args[i] < x) ++i;
    if (i > 1 && args[i - 1] > x) --i;
    const T* vals = args + n;
    const T* y2 = vals + n;
    const T h = args[i] - args[i - 1];
    const T aa = args[i] - x;
    const T bb = h - aa;
    const T cc = (aa + h) * y2[i - 1] + (bb + h) * y2[i];
    if (dS)
        *dS = ddiv(vals[i] - vals[i - 1] -
                   (T(1) / T(6)) * ((aa - bb) * cc + aa * bb * (y2[i] - y2[i - 1])), h);
    return ddiv(aa * vals[i - 1] + bb * vals[i] - (T(1) / T(6)) * aa * bb * cc, h);
}

// leaf_consts plus, for a Sersic leaf, the ends and the inverse spacing of its log-k knots in the
// arithmetic type (so the FP32 lookup needs neither those loads nor a division per sample).
template <typename T>
__device__ inline LeafC<T> leaf_consts_t(const gsmi_leaf& L, const Tables<T>& tb)
{
    LeafC<T> c = leaf_consts<T>(L);
    if (L.type == GSMI_LEAF_SERSIC && L.ntab > 1) {
        const T* args = sersic_table<T>(tb, L.table);
        c.t0 = args[0];
        c.t1 = args[L.ntab - 1];
        c.t2 = T(L.ntab - 1) / (c.t1 - c.t0);
    }
    return c;
}

// spline_eq in FP32 with the knot range and inverse spacing precomputed (leaf_consts_t).
__device__ inline float spline_eq32(const float* __restrict__ args, int n, float lk, float a0,
                                    float alast, float inv_da)
{
    if (lk < a0 || lk > alast) return 0.f;
    int i = int(ceilf((lk - a0) * inv_da));
    i = min(max(i, 1), n - 1);
    const float* vals = args + n;
    const float* y2 = vals + n;
    const float h = args[i] - args[i - 1];
    const float aa = args[i] - lk;
    const float bb = h - aa;
    const float cc = (aa + h) * y2[i - 1] + (bb + h) * y2[i];
    return ddiv(aa * vals[i - 1] + bb * vals[i] - (1.f / 6.f) * aa * bb * cc, h);
}

// spline_eq32 with the derivative dS/dlk (spline_eq's formula).
__device__ inline float spline_eq32_d(const float* __restrict__ args, int n, float lk, float a0,
                                      float alast, float inv_da, float* dS)
{
    if (lk < a0 || lk > alast) {
        *dS = 0.f;
        return 0.f;
    }
    int i = int(ceilf((lk - a0) * inv_da));
    i = min(max(i, 1), n - 1);
    const float* vals = args + n;
    const float* y2 = vals + n;
    const float h = args[i] - args[i - 1];
    const float ih = __builtin_amdgcn_rcpf(h);
    const float aa = args[i] - lk;
    const float bb = h - aa;
    const float cc = (aa + h) * y2[i - 1] + (bb + h) * y2[i];
    *dS = (vals[i] - vals[i - 1] - (1.f / 6.f) * ((aa - bb) * cc + aa * bb * (y2[i] - y2[i - 1])))
          * ih;
    return (aa * vals[i - 1] + bb * vals[i] - (1.f / 6.f) * aa * bb * cc) * ih;
}

// SersicInfo::kValue (src/SBSersic.cpp:363-376) in FP32 without branches: the high-k asymptote
// (a + b / k) / k^2 for ksq >= ksq_max, the Taylor series 1 + ksq (d2 + ksq d4) below ksq_min,
// and otherwise the table's natural cubic spline in ln k over ksq.  The spline is read as the
// per-interval coefficients c of S = c0 + s (c1 + s (c2 + s c3)), s = (ln k - knot i-1) / spacing
// in [0, 1] (the same cubic as spline_eq's form; the host derives them in FP64): one 16-byte
// load per sample.  The interval is Table::lookup's, i = ceil((ln k - a0) / spacing) clamped to
// [1, n-1] (src/Table.cpp:84-92), and the spline is 0 outside the knots; every index stays in
// [0, n-2] whatever ksq is (NaN included).
__device__ __forceinline__ float sersic_f32(const LeafC<float>& C, const float4* __restrict__ cf,
                                            int ntab, float ksq)
{
    const float r = __builtin_amdgcn_rsqf(ksq);
    const float asym = (C.p[2] + C.p[3] * r) * (r * r);
    const float tay = 1.f + ksq * (C.p[4] + ksq * C.p[5]);
    const float lk = 0.34657359027997264f * __builtin_amdgcn_logf(ksq);   // ln(ksq) / 2
    const float u = fminf(fmaxf((lk - C.t0) * C.t2, 0.f), float(ntab - 1));
    const int i = max(int(ceilf(u)), 1);
    const float s = u - float(i - 1);
    const float4 c = cf[i - 1];
    const float S = (lk < C.t0 || lk > C.t1) ? 0.f : c.x + s * (c.y + s * (c.z + s * c.w));
    const float spl = S * __builtin_amdgcn_rcpf(ksq);
    return ksq >= C.p[1] ? asym : (ksq < C.p[0] ? tay : spl);
}

// Exponential clipping: FP64 uses the row ranges the host computed as GetKValueRange1d/2d
// (src/SBProfile.cpp:508-588) for the stored k-image rows; FP32 clips at ksq > ksq_max.
template <typename T>
__device__ inline bool exp_clip(const LeafC<T>& C, const LeafI& I, const Tables<T>& tb, int i,
                                int j, int Nk, T ksq)
{
    if constexpr (sizeof(T) == 8) {
        if (I.aux >= 0) {
            const int32_t* r = tb.clip + I.aux + 2 * (j + Nk / 2);
            return i < r[0] || i >= r[1];
        }
    }
    return ksq > C.p[0];
}

// FP32 values of the radial leaves from ksq = |k'|^2 s^2: Exponential (clipped at ksq > p0, as
// exp_clip does in FP32), Kolmogorov, and Sersic read from its cubic coefficients (I.cf >= 0).
template <typename T>
__device__ __forceinline__ float radial_f32(const LeafC<T>& C, const LeafI& I,
                                            const Tables<T>& tb, float ksq)
{
    if (I.type == GSMI_LEAF_EXPONENTIAL) {
        if (ksq > C.p[0]) return 0.f;
        const float r = __builtin_amdgcn_rsqf(1.f + ksq);      // (1 + ksq)^(-3/2) = r^3
        return C.flux * r * r * r;
    }
    if (I.type == GSMI_LEAF_KOLMOGOROV) {
        const float t = __builtin_amdgcn_exp2f((5.f / 6.f) * __builtin_amdgcn_logf(ksq));
        return C.flux * __builtin_amdgcn_exp2f(-1.44269504088896341f * t);
    }
    return C.flux * sersic_f32(C, tb.coef + I.cf, I.ntab, ksq);
}

template <typename T>
__device__ inline void leaf_kprime(const LeafC<T>& C, const LeafI& I, T kx, T ky, T* kxp, T* kyp)
{
    if constexpr (sizeof(T) == 4) {
        // a separable leaf's record holds exact zeros off the diagonal, so the general form
        // gives the same products without a per-leaf select
        *kxp = C.j0 * kx + C.j2 * ky;
        *kyp = C.j1 * kx + C.j3 * ky;
        return;
    }
    if (I.sep) {
        *kxp = C.j0 * kx;
        *kyp = C.j3 * ky;
    } else {
        *kxp = C.j0 * kx + C.j2 * ky;
        *kyp = C.j1 * kx + C.j3 * ky;
    }
}

// Moffat of non-special beta in FP32: g(k) = k^nu K_nu(k), nu = beta - 1, and dg = g'(k)/(2k) =
// -k^(nu-1) K_{nu-1}(k) / 2, with the mu-only constants of Temme's series in C.p[2..5].  At k = 0:
// g = C.p[1] (the kV value at the origin, flux / knorm), dg = 0.
template <typename T>
__device__ inline void moffat_generic_f32(const LeafC<T>& C, float k, float* g, float* dg)
{
    if (k == 0.f) {
        *g = float(C.p[1]);
        *dg = 0.f;
        return;
    }
    const float nu = float(C.p[0]) - 1.f;
    const int nl = int(nu + 0.5f);
    const float mu = nu - float(nl);
    float km1;
    const float kn = bessel_k_f32(nl, mu, float(C.p[2]), float(C.p[3]), float(C.p[4]),
                                  float(C.p[5]), k, &km1);
    const float pk = __expf((nu - 1.f) * __logf(k));       // k^(nu - 1)
    *g = kn * pk * k;
    *dg = -0.5f * pk * km1;
}

// Airy k-space: the overlap area of two annuli (outer radius 1, inner radius obs) whose centres
// are sqrt(t) apart (AiryInfoObs::kValue and its helpers, src/SBAiry.cpp:259-329; the
// unobscured form AiryInfoNoObs::kValue, 409-420).
// (Clamped where rounding could take an argument just outside the domain; in the reference's
// FP64 arithmetic the clamps do not bind.)
template <typename T>
__device__ inline T airy_chord(T r, T h, T rsq, T hsq)
{
    if (r == T(0)) return T(0);
    return rsq * asin(fmin(ddiv(h, r), T(1))) - h * dsqrt(fmax(rsq - hsq, T(0)));
}

template <typename T>
__device__ inline T airy_circles(T r, T s, T rsq, T ssq, T tsq)     // radii r >= s
{
    const T rps_sq = (r + s) * (r + s);
    if (tsq >= rps_sq) return T(0);
    const T rms_sq = (r - s) * (r - s);
    if (tsq <= rms_sq) return T(M_PI) * ssq;
    const T hsq = fmax(T(0.5) * (rsq + ssq) - ddiv(tsq * tsq + rps_sq * rms_sq, T(4) * tsq),
                       T(0));
    const T h = dsqrt(hsq);
    if (tsq < rsq - ssq)
        return T(M_PI) * ssq - airy_chord(s, h, ssq, hsq) + airy_chord(r, h, rsq, hsq);
    return airy_chord(s, h, ssq, hsq) + airy_chord(r, h, rsq, hsq);
}

template <typename T>
__device__ inline T airy_circle_pair(T r, T rsq, T tsq)            // equal radii
{
    if (tsq >= T(4) * rsq) return T(0);
    if (tsq == T(0)) return T(M_PI) * rsq;
    const T hsq = rsq - tsq / T(4);
    return T(2) * airy_chord(r, dsqrt(hsq), rsq, hsq);
}

template <typename T>
__device__ inline T airy_kvalue(T t, T obs, T obssq)
{
    if (obs == T(0)) {
        if (t >= T(4)) return T(0);
        if (t == T(0)) return T(M_PI);
        const T hsq = T(1) - t / T(4);
        const T h = dsqrt(hsq);
        return T(2) * (asin(h) - h * dsqrt(T(1) - hsq));
    }
    return airy_circle_pair(T(1), T(1), t) - T(2) * airy_circles(T(1), obs, T(1), obssq, t) +
           airy_circle_pair(obs, obssq, t);
}

// Deconvolution._drawKImage's cut (galsim/convolve.py:602-612) at k-image pixel (x, y) = (i, j):
// k' = (J00 x + J10 y, J01 x + J11 y), cut where |k'|^2 dk^2 > maxk^2, in FP64 with the
// reference's order of operations (no fused multiply-adds) so the boundary samples agree.
__device__ inline bool kclip(const gsmi_leaf* __restrict__ rec, int i, int j)
{
    const double x = double(i), y = double(j);
    const double kxp = __dadd_rn(__dmul_rn(rec->jac[0], x), __dmul_rn(rec->jac[2], y));
    const double kyp = __dadd_rn(__dmul_rn(rec->jac[1], x), __dmul_rn(rec->jac[3], y));
    const double ksq = __dmul_rn(__dadd_rn(__dmul_rn(kxp, kxp), __dmul_rn(kyp, kyp)),
                                 __dmul_rn(rec->scale, rec->scale));
    return ksq > rec->p[0];
}

// 1 / z where |z| > a, else b: Deconvolution._drawKImage's inversion (galsim/convolve.py:
// 602-606), with numpy's complex reciprocal (ratio of the smaller to the larger component).
template <typename T>
__device__ inline cplx<T> deconv_inv(cplx<T> z, T a, T b)
{
    if (!(hypot(z.x, z.y) > a)) return {b, T(0)};
    if (fabs(z.x) >= fabs(z.y)) {
        const T rat = z.y / z.x;
        const T scl = T(1) / (z.x + z.y * rat);
        return {scl, -rat * scl};
    }
    const T rat = z.x / z.y;
    const T scl = T(1) / (z.y + z.x * rat);
    return {rat * scl, -scl};
}

// Value of a decoded leaf at k-image sample (i, j) (kx = i dk, ky = j dk).  rec is the leaf's
// record, read only by the FP64 Bessel-function Moffat kinds.
template <typename T>
__device__ inline T eval_leafc(const LeafC<T>& C, const LeafI& I, const gsmi_leaf* rec,
                               const Tables<T>& tb, T kx, T ky, int i, int j, int Nk)
{
    if constexpr (sizeof(T) == 4) {
        // the radial leaves need only |k'|^2 s^2, the leaf's quadratic form in (kx, ky): three
        // multiply-adds per leaf, with kx^2, kx ky and ky^2 shared by a program's leaves
        if (I.type == GSMI_LEAF_EXPONENTIAL || I.type == GSMI_LEAF_KOLMOGOROV ||
            (I.type == GSMI_LEAF_SERSIC && I.cf >= 0))
            return radial_f32(C, I, tb, C.q0 * (kx * kx) + C.q1 * (kx * ky) + C.q2 * (ky * ky));
    }
    T kxp, kyp;
    leaf_kprime(C, I, kx, ky, &kxp, &kyp);
    const T X = kxp * C.s;
    const T Y = kyp * C.s;
    const T flux = C.flux;
    switch (I.type) {
      case GSMI_LEAF_EXPONENTIAL: {
        // SBExponential::fillKImage (src/SBExponential.cpp:442-510)
        T ksq = X * X + Y * Y;
        if (exp_clip<T>(C, I, tb, i, j, Nk, ksq)) return T(0);
        T q = T(1) + ksq;
        if constexpr (sizeof(T) == 4) {
            const float r = __builtin_amdgcn_rsqf(q);         // q^(-3/2) = r^3
            return flux * r * r * r;
        }
        return ddiv(flux, q * dsqrt(q));
      }
      case GSMI_LEAF_GAUSSIAN: {
        // SBGaussian::fillKImage (src/SBGaussian.cpp:200-290)
        if (I.sep) return flux * dexp(T(-0.5) * X * X) * dexp(T(-0.5) * Y * Y);
        T ksq = X * X + Y * Y;
        if (ksq > C.p[0]) return T(0);
        if (ksq < C.p[1]) return flux * (T(1) - T(0.5) * ksq * (T(1) - T(0.25) * ksq));
        return flux * dexp(T(-0.5) * ksq);
      }
      case GSMI_LEAF_MOFFAT: {
        // SBMoffat kV_* (src/SBMoffat.cpp:286-345); flux holds knorm
        T ksq = X * X + Y * Y;
        T k = dsqrt(ksq);
        switch (I.kind) {
          case 1:
            return flux * moffat_trunc_spline<T>(sersic_table<T>(tb, I.aux), I.ntab, ksq, k,
                                                 C.p[1], nullptr);
          case 15: return flux * dexp(-k);
          case 25: return flux * dexp(-k) * (T(1) + k);
          case 35: return flux * dexp(-k) * (T(3) + (T(3) + k) * k);
          default:
            if constexpr (sizeof(T) == 4) {
                if (I.kind == 2 || I.kind == 3 || I.kind == 4) {
                    float g, dg;
                    moffat_int_kind_f32(I.kind, k, &g, &dg);
                    return flux * g;
                }
                if (I.kind == 0) {
                    float g, dg;
                    moffat_generic_f32(C, k, &g, &dg);
                    return flux * g;
                }
            }
            return flux * T(moffat_kv_bessel(I.kind, rec->p[0] - 1., double(k), double(ksq),
                                             rec->p[1]));
        }
      }
      case GSMI_LEAF_KOLMOGOROV: {
        // SBKolmogorov::fillKImage, KolmogorovInfo::kValue (src/SBKolmogorov.cpp:185-259)
        T ksq = X * X + Y * Y;
        if constexpr (sizeof(T) == 4) {
            // in base 2: ksq^(5/6) = 2^(5/6 log2 ksq), e^-t = 2^(-t log2 e); ksq = 0 gives 1
            const float t = __builtin_amdgcn_exp2f((5.f / 6.f) * __builtin_amdgcn_logf(ksq));
            return flux * __builtin_amdgcn_exp2f(-1.44269504088896341f * t);
        }
        return flux * dexp(-dexp(T(5. / 6.) * dlog(ksq)));
      }
      case GSMI_LEAF_SERSIC: {
        // SersicInfo::kValue (src/SBSersic.cpp:363-376)
        T ksq = X * X + Y * Y;
        if constexpr (sizeof(T) == 4) {
            if (I.cf >= 0) return flux * sersic_f32(C, tb.coef + I.cf, I.ntab, ksq);
            if (ksq >= C.p[1]) {
                const float r = __builtin_amdgcn_rsqf(ksq);
                return flux * (C.p[2] + C.p[3] * r) * (r * r);
            }
            if (ksq < C.p[0]) return flux * (1.f + ksq * (C.p[4] + ksq * C.p[5]));
            const float lk = 0.34657359027997264f * __builtin_amdgcn_logf(ksq);   // ln(ksq)/2
            return ddiv(flux * spline_eq32(sersic_table<float>(tb, I.aux), I.ntab, lk, C.t0,
                                           C.t1, C.t2), ksq);
        }
        if (ksq >= C.p[1]) return ddiv(flux * (C.p[2] + ddiv(C.p[3], dsqrt(ksq))), ksq);
        if (ksq < C.p[0]) return flux * (T(1) + ksq * (C.p[4] + ksq * C.p[5]));
        T lk = T(0.5) * dlog(ksq);
        return ddiv(flux * spline_eq<T>(sersic_table<T>(tb, I.aux), I.ntab, lk, nullptr), ksq);
      }
      case GSMI_LEAF_BOX: {
        // SBBox::fillKImage with math::sinc (src/SBBox.cpp:151-224, src/math/Sinc.cpp:29-33)
        const T Xb = kxp * C.s;
        const T Yb = kyp * C.s2;
        T sx = fabs(Xb) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * Xb * Xb
                                  : ddiv(dsinpi(Xb), T(M_PI) * Xb);
        T sy = fabs(Yb) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * Yb * Yb
                                  : ddiv(dsinpi(Yb), T(M_PI) * Yb);
        return flux * sx * sy;
      }
      case GSMI_LEAF_KCLIP:
        return kclip(rec, i, j) ? T(0) : T(1);
      case GSMI_LEAF_AIRY:
        // SBAiryImpl::fillKImage (src/SBAiry.cpp:189-248): knorm A((k lam/D / pi)^2)
        return flux * airy_kvalue<T>(X * X + Y * Y, C.p[0], C.p[1]);
      case GSMI_LEAF_INCLINED_EXP: {
        // SBInclinedExponentialImpl::kValueHelper (src/SBInclinedExponential.cpp:277-336):
        // a face-on Exponential at (kx, ky cos i) times the sech^2 disk's vertical factor
        const T kyc = Y * C.p[0];
        const T ksq = X * X + kyc * kyc;
        if (ksq > C.p[3]) return T(0);
        T base;
        if (ksq < C.p[2]) {
            base = T(1) - T(1.5) * ksq * (T(1) - T(1.25) * ksq);
        } else {
            const T t = T(1) + ksq;
            base = ddiv(T(1), t * dsqrt(t));
        }
        const T sky = C.p[1] * Y;
        const T sky2 = sky * sky;
        const T conv = sky2 < C.p[2] ? T(1) - T(0.16666666667) * sky2 *
                                                   (T(1) - T(0.116666666667) * sky2)
                                     : ddiv(sky, sinh(sky));
        return flux * base * conv;
      }
      case GSMI_LEAF_INCLINED_VERT: {
        // the clip and vertical factor of SBInclinedSersicImpl::kValueHelper
        // (src/SBInclinedSersic.cpp:321-370); the Sersic base is a separate leaf
        const T kyc = Y * C.p[0];
        if (X * X + kyc * kyc > C.p[3]) return T(0);
        const T sky = C.p[1] * Y;
        const T sky2 = sky * sky;
        return sky2 < C.p[2] ? T(1) - T(0.16666666667) * sky2 * (T(1) - T(0.116666666667) * sky2)
                             : ddiv(sky, sinh(sky));
      }
      case GSMI_LEAF_TOPHAT: {
        // SBTopHatImpl::kValue2 (src/SBBox.cpp:289-299)
        const T kr0sq = X * X + Y * Y;
        if (kr0sq < T(1.e-4)) return flux * (T(1) - kr0sq * (T(1. / 8.) + T(1. / 192.) * kr0sq));
        return flux * tophat_j1_ratio(dsqrt(kr0sq));
      }
      case GSMI_LEAF_DELTA:
        return flux;
      case GSMI_LEAF_SPERGEL: {
        // SBSpergelImpl::fillKImage, kloop_1d / kloop_2d (src/SBSpergel.cpp:121-139, 456-527)
        T ksq = X * X + Y * Y;
        if (exp_clip<T>(C, I, tb, i, j, Nk, ksq)) return T(0);
        if constexpr (sizeof(T) == 4)
            return flux * __builtin_amdgcn_exp2f(C.p[1] * __builtin_amdgcn_logf(1.f + ksq));
        return flux * pow(T(1) + ksq, C.p[1]);
      }
    }
    return T(0);
}

// Value of leaf record L at k-image sample (i, j), decoding it on the spot.
template <typename T>
__device__ inline T eval_leaf(const gsmi_leaf* __restrict__ Lp, const Tables<T>& tb, T kx, T ky,
                              int i, int j, int Nk)
{
    return eval_leafc<T>(leaf_consts_t<T>(*Lp, tb), leaf_ints(*Lp), Lp, tb, kx, ky, i, j, Nk);
}

// ---------------------------------------------------------------------------------------------
// InterpolatedImage (GSMI_LEAF_INTERP_IMAGE)
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ inline const T* table_at(const Tables<T>& tb, int off);
template <>
__device__ inline const double* table_at<double>(const Tables<double>& tb, int off)
{
    return tb.tab + off;
}
template <>
__device__ inline const float* table_at<float>(const Tables<float>& tb, int off)
{
    return tb.tab32 + off;
}

// math::sinc (src/math/Sinc.cpp:29-33)
template <typename T>
__device__ inline T sinc_pi(T x)
{
    if (fabs(x) < T(1e-4)) return T(1) - T(M_PI * M_PI / 6.) * x * x;
    return ddiv(dsinpi(x), T(M_PI) * x);
}

// Interpolant::xval (src/Interpolant.cpp:171-176, 245-250, 284-290, 389-433, 563-912): code as
// in the leaf's parameter block; n, dc and K(1..5) are Lanczos's.
template <typename T>
__device__ inline T interp_xval(int code, int n, bool dc, const double* __restrict__ K, T x)
{
    x = fabs(x);
    switch (code) {
      case 0: return x > T(0.5) ? T(0) : T(1);
      case 1: return x > T(1) ? T(0) : T(1) - x;
      case 2:
        if (x < T(1)) return T(1) + x * x * (T(1.5) * x - T(2.5));
        if (x < T(2)) return T(-0.5) * (x - T(1)) * (x - T(2)) * (x - T(2));
        return T(0);
      case 3:
        if (x <= T(1))
            return T(1) + x * x * x * (T(-95. / 12.) + x * (T(23. / 2.) + x * T(-55. / 12.)));
        if (x <= T(2))
            return (x - T(1)) * (x - T(2)) *
                   (T(-23. / 4.) + x * (T(29. / 2.) + x * (T(-83. / 8.) + x * T(55. / 24.))));
        if (x <= T(3))
            return (x - T(2)) * (x - T(3)) * (x - T(3)) *
                   (T(-9. / 4.) + x * (T(25. / 12.) + x * T(-11. / 24.)));
        return T(0);
      default: {
        const T nd = T(n);
        if (x >= nd) return T(0);
        T res, s;
        if (x > T(1e-4)) {
            s = dsinpi(x);
            res = ddiv(nd * T(1. / (M_PI * M_PI)) * s * dsinpi(ddiv(x, nd)), x * x);
        } else {
            const T pix = T(M_PI) * x;
            const T temp = T(1. / 6.) * pix * pix;
            s = pix * (T(1) - temp);
            res = T(1) - temp * (T(1) + ddiv(T(1), nd * nd));
        }
        if (dc) {
            const T ssq = s * s;
            const T factor =
                T(1) - T(4) * T(K[0]) * ssq - T(16) * T(K[1]) * ssq * (T(1) - ssq) -
                T(4) * T(K[2]) * ssq * (T(9) - ssq * (T(24) - T(16) * ssq)) -
                T(64) * T(K[3]) * ssq * (T(1) - ssq * (T(5) - ssq * (T(8) - T(4) * ssq))) -
                T(4) * T(K[4]) * ssq *
                    (T(25) - ssq * (T(200) - ssq * (T(560) - ssq * (T(640) - T(256) * ssq))));
            res = ddiv(res, factor);
        }
        return res;
      }
    }
}

// Interpolant::uval (src/Interpolant.cpp:176, 251-255, 292-312, 435-462, 914-919); Lanczos reads
// its spline table (knots, values, second derivatives; 0 beyond umax).
template <typename T>
__device__ inline T interp_uval(int code, const T* __restrict__ utab, int ntab, T umax, T u)
{
    u = fabs(u);
    switch (code) {
      case 0: return sinc_pi(u);
      case 1: {
        const T s = sinc_pi(u);
        return s * s;
      }
      case 2: {
        const T s = sinc_pi(u);
        const T c = cos(T(M_PI) * u);
        return s * s * s * (T(3) * s - T(2) * c);
      }
      case 3: {
        const T s = sinc_pi(u);
        const T piu = T(M_PI) * u;
        const T c = cos(piu);
        const T ssq = s * s;
        const T piusq = piu * piu;
        return s * ssq * ssq * (s * (T(55) - T(19) * piusq) + T(2) * c * (piusq - T(27)));
      }
      default:
        return u > umax ? T(0) : spline_eq<T>(utab, ntab, u, nullptr);
    }
}

// Taps of the k interpolation at scaled k: ceil(k - xr) .. floor(k + xr), or the one at an
// integer k when the form checks for it (src/SBInterpolatedImage.cpp:415-427, 827-840).
template <typename T>
__device__ inline void interp_taps(T k, T xr, bool icheck, int* first, int* count)
{
    const T fl = floor(k + T(0.01));
    if (icheck && fabs(k - fl) < T(10. * 2.220446049250313e-16) * (fabs(k) + T(1))) {
        *first = int(fl);
        *count = 1;
        return;
    }
    *first = int(ceil(k - xr));
    *count = min(int(floor(k + xr)) - *first + 1, GSMI_INTERP_MAX_TAPS);
}

// WrapKIndex (src/SBInterpolatedImage.cpp:217-222)
__device__ inline int wrap_k_index(int k, int No2, int N)
{
    k = (k + No2) % N;
    if (k < 0) k += N;
    return k - No2;
}

// The leaf's complex value at k-image sample (i, j) (kx = i dk, ky = j dk): k in the image's
// frame through the leaf's Jacobian, the cut of its form, then sum_q ywt(q) sum_p xwt(p) K(p, q)
// over the wrapped taps as KValueInnerLoop walks them (p runs N/2 + 1 -> -N/2 + 1, q runs
// N/2 -> -N/2; p < 0 reads the conjugate at (-p, -q), q = -N/2 its own row), times
// uval(kx / 2 pi) uval(ky / 2 pi) of the x-interpolant.
template <typename T>
__device__ inline cplx<T> interp_image_value(const gsmi_leaf* __restrict__ L, const Tables<T>& tb,
                                             T kx, T ky, int i, int j)
{
    const double* pb = tb.tab + L->table;
    const bool kv = L->kind == 1;
    const bool sep = L->separable != 0;
    T kxp, kyp;
    if (sep) {
        kxp = T(L->jac[0]) * kx;
        kyp = T(L->jac[3]) * ky;
    } else {
        kxp = T(L->jac[0]) * kx + T(L->jac[2]) * ky;
        kyp = T(L->jac[1]) * kx + T(L->jac[3]) * ky;
    }
    const cplx<T> zero = {T(0), T(0)};
    if (kv) {
        const T m1 = T(pb[GSMI_INTERP_MAXK1]);
        if (fabs(kxp) > m1 || fabs(kyp) > m1) return zero;
    } else if (sep) {
        const int ir = i - int(pb[GSMI_INTERP_XMIN]);
        const int jr = j - int(pb[GSMI_INTERP_YMIN]);
        if (ir < int(pb[GSMI_INTERP_I1]) || ir >= int(pb[GSMI_INTERP_I2]) ||
            jr < int(pb[GSMI_INTERP_J1]) || jr >= int(pb[GSMI_INTERP_J2]))
            return zero;
    }
    const int No2 = int(pb[GSMI_INTERP_N2]);
    const int N = 2 * No2;
    const T kscale = T(pb[GSMI_INTERP_KSCALE]);
    const T kxs = kxp * kscale;
    const T kys = kyp * kscale;
    if (!kv && !sep) {
        const T m1s = T(pb[GSMI_INTERP_MAXK1S]);
        if (fabs(kxs) > m1s || fabs(kys) > m1s) return zero;
    }
    const bool icheck = kv || sep;
    const T xr = T(pb[GSMI_INTERP_KXR]);
    int p1, np, q1, nq;
    interp_taps(kxs, xr, icheck, &p1, &np);
    interp_taps(kys, xr, icheck, &q1, &nq);
    const int kcode = int(pb[GSMI_INTERP_KCODE]);
    const int kn = int(pb[GSMI_INTERP_KN]);
    const bool kdc = pb[GSMI_INTERP_KDC] != 0.;
    const double* KK = pb + GSMI_INTERP_KK;
    const T* K = table_at<T>(tb, int(pb[GSMI_INTERP_KIMAGE]));
    const int W = No2 + 1;
    const int pw0 = wrap_k_index(p1, No2, N);
    int qw = wrap_k_index(q1, No2, N);
    cplx<T> sum = zero;
    for (int t = 0; t < nq; ++t, ++qw) {
        if (qw == No2) qw -= N;
        const T ywt = interp_xval<T>(kcode, kn, kdc, KK, T(q1 + t) - kys);
        const int mq = qw == -No2 ? qw : -qw;
        cplx<T> xs = zero;
        int pw = pw0;
        for (int s = 0; s < np; ++s, ++pw) {
            if (pw == No2 + 1) pw -= N;
            const T xwt = interp_xval<T>(kcode, kn, kdc, KK, T(p1 + s) - kxs);
            const bool neg = pw < 0;
            const T* e = K + 2 * (size_t((neg ? mq : qw) + No2) * W + (neg ? -pw : pw));
            xs.x += xwt * e[0];
            xs.y += xwt * (neg ? -e[1] : e[1]);
        }
        sum.x += ywt * xs.x;
        sum.y += ywt * xs.y;
    }
    const int xcode = int(pb[GSMI_INTERP_XCODE]);
    const int xt = int(pb[GSMI_INTERP_XUTAB]);
    const T* utab = xt >= 0 ? table_at<T>(tb, xt) : nullptr;
    const int un = int(pb[GSMI_INTERP_XUTABN]);
    const T umax = T(pb[GSMI_INTERP_XUMAX]);
    const T f = interp_uval<T>(xcode, utab, un, umax, kxp * T(1. / (2. * M_PI))) *
                interp_uval<T>(xcode, utab, un, umax, kyp * T(1. / (2. * M_PI)));
    return {f * sum.x, f * sum.y};
}

// A small stack held in registers: push/pop shift every slot, so all indices are static.
template <typename T, int S>
struct RegStack {
    cplx<T> s[S];
    __device__ __forceinline__ void push(cplx<T> v)
    {
#pragma unroll
        for (int k = S - 1; k > 0; --k) s[k] = s[k - 1];
        s[0] = v;
    }
    __device__ __forceinline__ cplx<T> pop()
    {
        cplx<T> v = s[0];
#pragma unroll
        for (int k = 0; k < S - 1; ++k) s[k] = s[k + 1];
        return v;
    }
};

template <typename T>
struct StackDepth;
template <>
struct StackDepth<double> { static constexpr int value = GSMI_STACK_F64; };
template <>
struct StackDepth<float> { static constexpr int value = GSMI_STACK_F32; };

// A stamp program decoded once per workgroup: op codes and arguments packed in uniform
// registers (code | arg << 8), leaf integer fields in uniform registers, leaf constants in LDS.
// Programs with more leaves or ops than this run on the record-reading interpreter.
constexpr int kProgLeaves = 4;
constexpr int kProgOps = 12;

struct ProgramD {
    int nl, nops;
    bool has_phase;       // an in-program PHASE: the program value is complex
    int opc[kProgOps];
    LeafI li[kProgLeaves];
};

// Number of leaf slots a program uses (1 + the largest LEAF argument).
__device__ inline int program_leaves(const gsmi_op* __restrict__ ops, int nops)
{
    int nl = 0;
    for (int o = 0; o < nops; ++o)
        if (ops[o].code == GSMI_OP_LEAF) nl = max(nl, ops[o].arg + 1);
    return nl;
}

// Decode the program (all threads) and write its leaf constants to lc[] (threads < nl).  The
// caller synchronises before reading lc.  Returns false if the program does not fit.
template <typename T>
__device__ inline bool decode_program(const gsmi_op* __restrict__ ops, int nops,
                                      const gsmi_leaf* __restrict__ leaves, const Tables<T>& tb,
                                      ProgramD& d, LeafC<T>* lc)
{
    const int nl = program_leaves(ops, nops);
    d.nl = nl;
    d.nops = nops;
    d.has_phase = false;
    if (nl > kProgLeaves || nops > kProgOps) return false;
    for (int o = 0; o < nops; ++o)
        if (ops[o].code == GSMI_OP_PHASE) d.has_phase = true;
#pragma unroll
    for (int o = 0; o < kProgOps; ++o)
        d.opc[o] = (o < nops) ? (ops[o].code | (ops[o].arg << 8)) : 0;
#pragma unroll
    for (int l = 0; l < kProgLeaves; ++l)
        d.li[l] = (l < nl) ? leaf_ints(leaves[l]) : LeafI{0, 0, 0, -1, 0, -1};
    if (int(threadIdx.x) < nl) lc[threadIdx.x] = leaf_consts_t<T>(leaves[threadIdx.x], tb);
    return true;
}

// The decoded program on leaf values v[] (LEAF ops take v[arg]; slot ovl, if >= 0, replaced by
// ovv).  SCALE and PHASE read their constants from the op records.
template <typename T>
__device__ inline cplx<T> run_ops(const ProgramD& d, const gsmi_op* __restrict__ ops,
                                  const T (&v)[kProgLeaves], int ovl, T ovv, T kx, T ky)
{
    RegStack<T, StackDepth<T>::value> st;
#pragma unroll
    for (int o = 0; o < kProgOps; ++o) {
        if (o < d.nops) {
            const int code = d.opc[o] & 0xff;
            const int a = d.opc[o] >> 8;
            switch (code) {
              case GSMI_OP_LEAF: {
                T x = v[0];
#pragma unroll
                for (int l = 1; l < kProgLeaves; ++l)
                    if (a == l) x = v[l];
                if (a == ovl) x = ovv;
                st.push({x, T(0)});
                break;
              }
              case GSMI_OP_MUL: {
                cplx<T> b = st.pop();
                st.s[0] = cmul(st.s[0], b);
                break;
              }
              case GSMI_OP_ADD: {
                cplx<T> b = st.pop();
                st.s[0] = cadd(st.s[0], b);
                break;
              }
              case GSMI_OP_SCALE:
                st.s[0].x *= T(ops[o].a);
                st.s[0].y *= T(ops[o].a);
                break;
              case GSMI_OP_PHASE: {
                T arg = T(ops[o].a) * kx + T(ops[o].b) * ky;
                T sn, cs;
                dsincos(arg, &sn, &cs);
                const T fs = T(ops[o].c);
                st.s[0] = cmul(st.s[0], cplx<T>{fs * cs, -fs * sn});
                break;
              }
              case GSMI_OP_INV:
                st.s[0] = deconv_inv(st.s[0], T(ops[o].a), T(ops[o].b));
                break;
            }
        }
    }
    return st.s[0];
}

// The same for a program without PHASE ops, whose value is real (radial leaves are real in
// k-space): half the registers of the complex form.
template <typename T>
__device__ inline T run_ops_real(const ProgramD& d, const gsmi_op* __restrict__ ops,
                                 const T (&v)[kProgLeaves], int ovl, T ovv)
{
    T st[StackDepth<T>::value];
#pragma unroll
    for (int k = 0; k < StackDepth<T>::value; ++k) st[k] = T(0);
#pragma unroll
    for (int o = 0; o < kProgOps; ++o) {
        if (o < d.nops) {
            const int code = d.opc[o] & 0xff;
            const int a = d.opc[o] >> 8;
            if (code == GSMI_OP_LEAF) {
                T x = v[0];
#pragma unroll
                for (int l = 1; l < kProgLeaves; ++l)
                    if (a == l) x = v[l];
                if (a == ovl) x = ovv;
#pragma unroll
                for (int k = StackDepth<T>::value - 1; k > 0; --k) st[k] = st[k - 1];
                st[0] = x;
            } else if (code == GSMI_OP_MUL || code == GSMI_OP_ADD) {
                const T b = st[0];
#pragma unroll
                for (int k = 0; k < StackDepth<T>::value - 1; ++k) st[k] = st[k + 1];
                st[0] = (code == GSMI_OP_MUL) ? st[0] * b : st[0] + b;
            } else if (code == GSMI_OP_SCALE) {
                st[0] *= T(ops[o].a);
            } else if (code == GSMI_OP_INV) {
                st[0] = fabs(st[0]) > T(ops[o].a) ? T(1) / st[0] : T(ops[o].b);
            }
        }
    }
    return st[0];
}

// Program shapes compiled into specialised fold kernels: the op sequence and the leaf types are
// compile-time constants there, so the sample loop has no interpreter (no op fetch or dispatch)
// and every leaf branch but the right one folds away.  The kernels check each stamp against the
// shape and take the interpreter for any stamp that does not match.
struct StaticProgram {
    int nops;
    int code[5];
    int arg[5];
    int nl;
    int type[3];
};

// index 0 is unused (the interpreter); the table's order is the gsmi_schedule program id
constexpr StaticProgram kStaticPrograms[] = {
    {0, {0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, 0, {0, 0, 0}},
    // (Exponential + Sersic) x Kolmogorov: a bulge+disk galaxy under atmospheric seeing
    {5, {GSMI_OP_LEAF, GSMI_OP_LEAF, GSMI_OP_ADD, GSMI_OP_LEAF, GSMI_OP_MUL}, {0, 1, 0, 2, 0}, 3,
     {GSMI_LEAF_EXPONENTIAL, GSMI_LEAF_SERSIC, GSMI_LEAF_KOLMOGOROV}},
    // Exponential x Moffat
    {3, {GSMI_OP_LEAF, GSMI_OP_LEAF, GSMI_OP_MUL, 0, 0}, {0, 1, 0, 0, 0}, 2,
     {GSMI_LEAF_EXPONENTIAL, GSMI_LEAF_MOFFAT, 0}},
    // Sersic x Kolmogorov
    {3, {GSMI_OP_LEAF, GSMI_OP_LEAF, GSMI_OP_MUL, 0, 0}, {0, 1, 0, 0, 0}, 2,
     {GSMI_LEAF_SERSIC, GSMI_LEAF_KOLMOGOROV, 0}},
    // Exponential x Kolmogorov
    {3, {GSMI_OP_LEAF, GSMI_OP_LEAF, GSMI_OP_MUL, 0, 0}, {0, 1, 0, 0, 0}, 2,
     {GSMI_LEAF_EXPONENTIAL, GSMI_LEAF_KOLMOGOROV, 0}},
};
constexpr int kNumStaticPrograms = 5;

// The Sersic leaf slot of program shape SIG, or -1.
template <int SIG>
constexpr int sersic_slot()
{
    for (int l = 0; l < kStaticPrograms[SIG].nl; ++l)
        if (kStaticPrograms[SIG].type[l] == GSMI_LEAF_SERSIC) return l;
    return -1;
}

// Does the stamp's program (ops, leaves) have shape SIG?
template <int SIG>
__device__ inline bool matches_static(const gsmi_op* __restrict__ ops, int nops,
                                      const gsmi_leaf* __restrict__ leaves)
{
    constexpr StaticProgram P = kStaticPrograms[SIG];
    if (nops != P.nops) return false;
#pragma unroll
    for (int o = 0; o < P.nops; ++o)
        if (ops[o].code != P.code[o] || ops[o].arg != P.arg[o]) return false;
#pragma unroll
    for (int l = 0; l < P.nl; ++l)
        if (leaves[l].type != P.type[l]) return false;
    return true;
}

// The ops of program shape SIG on leaf values v[] (slot ovl, if >= 0, replaced by ovv); with
// constant arguments the stack and the selects fold into straight-line arithmetic.
template <typename T, int SIG>
__device__ __forceinline__ T static_ops(const T (&v)[kProgLeaves], int ovl, T ovv)
{
    constexpr StaticProgram P = kStaticPrograms[SIG];
    T st[3] = {};
    int sp = 0;
#pragma unroll
    for (int o = 0; o < P.nops; ++o) {
        if (P.code[o] == GSMI_OP_LEAF) {
            st[sp++] = (P.arg[o] == ovl) ? ovv : v[P.arg[o]];
        } else if (P.code[o] == GSMI_OP_MUL) {
            --sp;
            st[sp - 1] = st[sp - 1] * st[sp];
        } else if (P.code[o] == GSMI_OP_ADD) {
            --sp;
            st[sp - 1] = st[sp - 1] + st[sp];
        }
    }
    return st[0];
}

// The value of program shape SIG at sample (i, j), its leaf constants in lc[] (LDS) and integer
// fields in li[] (their type overridden by the compile-time one).
template <typename T, int SIG>
__device__ __forceinline__ T eval_static(const LeafC<T>* __restrict__ lc,
                                         const LeafI (&li)[kProgLeaves],
                                         const gsmi_leaf* __restrict__ leaves,
                                         const Tables<T>& tb, T kx, T ky, int i, int j, int Nk)
{
    constexpr StaticProgram P = kStaticPrograms[SIG];
    T v[kProgLeaves] = {};
#pragma unroll
    for (int l = 0; l < P.nl; ++l) {
        LeafI I = li[l];
        I.type = P.type[l];
        if constexpr (sizeof(T) == 4) {
            if (P.type[l] == GSMI_LEAF_SERSIC) {
                // specialised FP32 kernels stage the coefficients at tb.coef (LeafI::cf = 0)
                const T ksq = lc[l].q0 * (kx * kx) + lc[l].q1 * (kx * ky) + lc[l].q2 * (ky * ky);
                v[l] = lc[l].flux * sersic_f32(lc[l], tb.coef, I.ntab, ksq);
                continue;
            }
        }
        v[l] = eval_leafc<T>(lc[l], I, leaves + l, tb, kx, ky, i, j, Nk);
    }
    return static_ops<T, SIG>(v, -1, T(0));
}

// Leaf values of a decoded program at sample (i, j).
template <typename T>
__device__ inline void eval_leaves(const ProgramD& d, const LeafC<T>* __restrict__ lc,
                                   const gsmi_leaf* __restrict__ leaves, const Tables<T>& tb,
                                   T kx, T ky, int i, int j, int Nk, T (&v)[kProgLeaves])
{
#pragma unroll
    for (int l = 0; l < kProgLeaves; ++l)
        v[l] = (l < d.nl) ? eval_leafc<T>(lc[l], d.li[l], leaves + l, tb, kx, ky, i, j, Nk)
                          : T(0);
}

// Evaluate the stamp program at k-image sample (i, j).
template <typename T>
__device__ inline cplx<T> eval_program(const gsmi_op* __restrict__ ops, int nops,
                                       const gsmi_leaf* __restrict__ leaves, const Tables<T>& tb,
                                       T dk, int i, int j, int Nk)
{
    RegStack<T, StackDepth<T>::value> st;
    const T kx = T(i) * dk;
    const T ky = T(j) * dk;
    for (int o = 0; o < nops; ++o) {
        const gsmi_op* op = ops + o;
        switch (op->code) {
          case GSMI_OP_LEAF: {
            const gsmi_leaf* L = leaves + op->arg;
            if (L->type == GSMI_LEAF_INTERP_IMAGE) {
                st.push(interp_image_value<T>(L, tb, kx, ky, i, j));
                break;
            }
            T v = eval_leaf<T>(L, tb, kx, ky, i, j, Nk);
            st.push({v, T(0)});
            break;
          }
          case GSMI_OP_MUL: {
            cplx<T> b = st.pop();
            st.s[0] = cmul(st.s[0], b);
            break;
          }
          case GSMI_OP_ADD: {
            cplx<T> b = st.pop();
            st.s[0] = cadd(st.s[0], b);
            break;
          }
          case GSMI_OP_SCALE:
            st.s[0].x *= T(op->a);
            st.s[0].y *= T(op->a);
            break;
          case GSMI_OP_PHASE: {
            T arg = T(op->a) * kx + T(op->b) * ky;
            T sn, cs;
            dsincos(arg, &sn, &cs);
            const T fs = T(op->c);
            st.s[0] = cmul(st.s[0], cplx<T>{fs * cs, -fs * sn});
            break;
          }
          case GSMI_OP_INV:
            st.s[0] = deconv_inv(st.s[0], T(op->a), T(op->b));
            break;
        }
    }
    return st.s[0];
}

// Separable factors hoisted out of the program (pixel sinc with a diagonal Jacobian, the phase
// of a shift, constant scalings): X(kx) for the stored kx = 0..Nk/2 and Y(ky) for ky =
// -Nk/2..Nk/2.  Constants and flux go into X.
template <typename T>
__device__ inline cplx<T> axis_factor(const gsmi_op* __restrict__ ops, int nops,
                                      const gsmi_leaf* __restrict__ leaves, T k, bool xaxis)
{
    cplx<T> f = {T(1), T(0)};
    for (int o = 0; o < nops; ++o) {
        const gsmi_op* op = ops + o;
        switch (op->code) {
          case GSMI_OP_LEAF: {
            const gsmi_leaf* L = leaves + op->arg;     // a Box with a diagonal Jacobian
            const T u = xaxis ? T(L->jac[0]) * k * T(L->scale) : T(L->jac[3]) * k * T(L->scale2);
            T sc = fabs(u) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * u * u
                                     : dsinpi(u) / (T(M_PI) * u);
            if (xaxis) sc *= T(L->flux);
            f.x *= sc;
            f.y *= sc;
            break;
          }
          case GSMI_OP_SCALE:
            if (xaxis) { f.x *= T(op->a); f.y *= T(op->a); }
            break;
          case GSMI_OP_PHASE: {
            T arg = (xaxis ? T(op->a) : T(op->b)) * k;
            T sn, cs;
            dsincos(arg, &sn, &cs);
            const T fs = xaxis ? T(op->c) : T(1);
            f = cmul(f, cplx<T>{fs * cs, -fs * sn});
            break;
          }
        }
    }
    return f;
}

// Twiddle tables exp(+2 pi i m / P) for the largest supported sizes of each family: a size-N
// transform reads every (P / N)-th entry of the table of its family (2^k or 3 * 2^k).
constexpr int kTwPow2 = 4096;
constexpr int kTwPow3 = 3072;

template <typename T>
struct TwiddleSet {
    const cplx<T>* p2;
    const cplx<T>* p3;
    __device__ __forceinline__ const cplx<T>* pick(int N, int* stride) const
    {
        if (N % 3 == 0) {
            *stride = kTwPow3 / N;
            return p3;
        }
        *stride = kTwPow2 / N;
        return p2;
    }
};

// ---------------------------------------------------------------------------------------------
// In-LDS Stockham FFT (inverse, e^{+2 pi i}) over nseg contiguous segments of length n.
// tw[m * twstride] = exp(+2 pi i m / n).  Mixed radix 4/2/3; returns the buffer with the result.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ inline void bfly2(cplx<T>* v)
{
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T>
__device__ inline void bfly4(cplx<T>* v)
{
    cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cplx<T> a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    cplx<T> a3 = {-d.y, d.x};   // i * (x1 - x3)
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

// Radix 8 as two radix-4 transforms of the even and odd samples joined by the eighth roots of
// unity w^k = e^{+2 pi i k / 8}.
template <typename T>
__device__ inline void bfly8(cplx<T>* v)
{
    const T h = T(0.70710678118654752440);   // 1/sqrt(2)
    cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
    bfly4(e);
    bfly4(o);
    const cplx<T> t1 = {h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)};    // w   o1
    const cplx<T> t2 = {-o[2].y, o[2].x};                                 // w^2 o2 = i o2
    const cplx<T> t3 = {-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)};   // w^3 o3
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
}

template <typename T>
__device__ inline void bfly3(cplx<T>* v)
{
    const T h = T(0.86602540378443864676);   // sqrt(3)/2
    cplx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cplx<T> m = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
    cplx<T> id = {-h * d.y, h * d.x};        // i sqrt(3)/2 (x1 - x2)
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, id);
    v[2] = csub(m, id);
}

// Radix 5 with w = e^{+2 pi i / 5}: Y1,4 = x0 + c1 (x1 + x4) + c2 (x2 + x3) +- i (s1 (x1 - x4) +
// s2 (x2 - x3)), Y2,3 = x0 + c2 (x1 + x4) + c1 (x2 + x3) +- i (s2 (x1 - x4) - s1 (x2 - x3)).
template <typename T>
__device__ inline void bfly5(cplx<T>* v)
{
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);   // cos 2pi/5, 4pi/5
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);    // sin 2pi/5, 4pi/5
    const cplx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const cplx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const cplx<T> a1 = {v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
    const cplx<T> a2 = {v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
    const cplx<T> b1 = {s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y};
    const cplx<T> b2 = {s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y};
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = {a1.x - b1.y, a1.y + b1.x};
    v[4] = {a1.x + b1.y, a1.y - b1.x};
    v[2] = {a2.x - b2.y, a2.y + b2.x};
    v[3] = {a2.x + b2.y, a2.y - b2.x};
}

// Radix 12 (e^{+2 pi i}) as radix 3 over n1 and radix 4 over n2 with n = 4 n1 + n2, k = k1 + 3 k2:
// w12^{nk} = w3^{n1 k1} w12^{n2 k1} w4^{n2 k2}.
template <typename T>
__device__ inline void bfly12(cplx<T>* v)
{
    const T h = T(0.86602540378443864676);   // sqrt(3)/2
    cplx<T> a[4][3];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cplx<T> u[3] = {v[n2], v[n2 + 4], v[n2 + 8]};
        bfly3(u);
        a[n2][0] = u[0];
        a[n2][1] = u[1];
        a[n2][2] = u[2];
    }
    // w12^m, m = n2 k1: w^1 = (h, 1/2), w^2 = (1/2, h), w^3 = i, w^4 = (-1/2, h), w^6 = -1
    a[1][1] = cmul(a[1][1], cplx<T>{h, T(0.5)});
    a[1][2] = cmul(a[1][2], cplx<T>{T(0.5), h});
    a[2][1] = cmul(a[2][1], cplx<T>{T(0.5), h});
    a[2][2] = cmul(a[2][2], cplx<T>{T(-0.5), h});
    a[3][1] = {-a[3][1].y, a[3][1].x};
    a[3][2] = {-a[3][2].x, -a[3][2].y};
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        cplx<T> u[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        bfly4(u);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 3 * k2] = u[k2];
    }
}

// LDS layout of the padded transforms: element i of a line at lds_pad(i), one slot of padding
// after every 32 elements (256 bytes of FP32 complex values, one sweep of the 64 four-byte
// banks), lines lds_line(n) apart.  The Stockham writes of stride R (and the multi-line reads)
// then fall on different banks instead of piling onto a few.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int lds_line(int n) { return n + (n >> 5); }

template <bool PAD>
__device__ __forceinline__ int lpos(int i)
{
    if constexpr (PAD) return lds_pad(i);
    return i;
}

template <int N>
constexpr bool kPow2 = N > 0 && (N & (N - 1)) == 0;

// Padded offset of element b + r S from element b, for powers of two S and b a multiple of
// min(S, 32) plus less than S (so b + r S crosses 32-runs exactly where r S does):
// r S + (r S) / 32.
template <int S>
__device__ __forceinline__ constexpr int pad_step(int r)
{
    return r * S + ((r * S) >> 5);
}

template <typename T, int R, bool PAD = false>
__device__ __forceinline__ void stockham_pass(const cplx<T>* src, cplx<T>* dst, int n, int nseg, int Ns,
                                     const cplx<T>* tw, int twstride)
{
    const int line = PAD ? lds_line(n) : n;
    const int nb = n / R;
    const int tstep = (n / (Ns * R)) * twstride;
    const float inv_nb = 1.f / float(nb);
    const float inv_ns = 1.f / float(Ns);
    for (int t = threadIdx.x; t < nseg * nb; t += blockDim.x) {
        const int seg = fdiv(t, inv_nb);
        const int tt = t - seg * nb;
        const int tq = fdiv(tt, inv_ns);
        const int jj = tt - tq * Ns;
        const cplx<T>* s = src + seg * line;
        cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[lpos<PAD>(tt + r * nb)];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[jj * r * tstep]);
        }
        if (R == 2) bfly2(v);
        if (R == 3) bfly3(v);
        if (R == 4) bfly4(v);
        if (R == 5) bfly5(v);
        if (R == 8) bfly8(v);
        cplx<T>* d = dst + seg * line;
        const int d0 = tq * Ns * R + jj;
#pragma unroll
        for (int r = 0; r < R; ++r) d[lpos<PAD>(d0 + r * Ns)] = v[r];
    }
}

template <typename T, bool PAD = false>
__device__ __forceinline__ cplx<T>* fft_lds(cplx<T>* a, cplx<T>* b, int n, int nseg, const cplx<T>* tw,
                            int twstride)
{
    int Ns = 1;
    int rem = n;
    cplx<T>* src = a;
    cplx<T>* dst = b;
    while (rem > 1) {
        if (rem % 3 == 0) {
            stockham_pass<T, 3, PAD>(src, dst, n, nseg, Ns, tw, twstride);
            Ns *= 3; rem /= 3;
        } else if (rem % 4 == 0) {
            stockham_pass<T, 4, PAD>(src, dst, n, nseg, Ns, tw, twstride);
            Ns *= 4; rem /= 4;
        } else {
            stockham_pass<T, 2, PAD>(src, dst, n, nseg, Ns, tw, twstride);
            Ns *= 2; rem /= 2;
        }
        __syncthreads();
        cplx<T>* tmp = src; src = dst; dst = tmp;
    }
    return src;
}

// The same transform with the FFT length a compile-time constant: the flat-index divisions of
// each pass become multiplications by constants, the twiddles of a butterfly are loaded together
// before they are used, and radix-8 passes (radix 3 first, then 8, 4, 2) cut the number of LDS
// round trips (128 = 8 8 2, 192 = 3 8 8, 256 = 8 8 4, 512 = 8 8 8).
template <typename T, int N, int R, int NS, bool PAD = false>
__device__ __forceinline__ void stockham_pass_c(const cplx<T>* src, cplx<T>* dst, int nseg,
                                                const cplx<T>* tw, int twstride)
{
    constexpr unsigned line = PAD ? lds_line(N) : N;
    constexpr unsigned nb = N / R;
    const int tstep = (N / (NS * R)) * twstride;
    const unsigned total = unsigned(nseg) * nb;
    for (unsigned t = threadIdx.x; t < total; t += blockDim.x) {
        const unsigned seg = t / nb;
        const unsigned tt = t - seg * nb;
        const unsigned tq = tt / NS;
        const unsigned jj = tt - tq * NS;
        const cplx<T>* s = src + seg * line;
        cplx<T> v[R];
        if constexpr (PAD && kPow2<N>) {
            // tt < nb and both powers of two: element tt + r nb sits at lds_pad(tt) plus a
            // compile-time offset (a multiple of 32 when nb >= 32, within one 32-run otherwise)
            const cplx<T>* s0 = s + lds_pad(int(tt));
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = s0[pad_step<nb>(r)];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = s[lpos<PAD>(int(tt + r * nb))];
        }
        if constexpr (NS > 1) {
            cplx<T> w[R];
#pragma unroll
            for (int r = 1; r < R; ++r) w[r] = tw[int(jj) * r * tstep];
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[r]);
        }
        if constexpr (R == 2) bfly2(v);
        if constexpr (R == 3) bfly3(v);
        if constexpr (R == 4) bfly4(v);
        if constexpr (R == 8) bfly8(v);
        cplx<T>* d = dst + seg * line;
        const unsigned d0 = tq * (NS * R) + jj;
        if constexpr (PAD && kPow2<N>) {
            // d0 = tq NS R + jj, jj < NS: element d0 + r NS sits at lds_pad(d0) plus
            // pad_step<NS>(r) (the block tq NS R never straddles a 32-run differently per r)
            cplx<T>* d1 = d + lds_pad(int(d0));
#pragma unroll
            for (int r = 0; r < R; ++r) d1[pad_step<NS>(r)] = v[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) d[lpos<PAD>(int(d0 + r * NS))] = v[r];
        }
    }
}

template <typename T, int N, int NS, bool PAD = false>
__device__ __forceinline__ cplx<T>* fft_fixed(cplx<T>* src, cplx<T>* dst, int nseg,
                                              const cplx<T>* tw, int twstride)
{
    constexpr int rem = N / NS;
    if constexpr (rem == 1) {
        return src;
    } else {
        constexpr int R = (rem % 3 == 0) ? 3 : (rem % 8 == 0) ? 8 : (rem % 4 == 0) ? 4 : 2;
        stockham_pass_c<T, N, R, NS, PAD>(src, dst, nseg, tw, twstride);
        __syncthreads();
        return fft_fixed<T, N, NS * R, PAD>(dst, src, nseg, tw, twstride);
    }
}

// fft_lds with the common lengths (64 to 3072, the sizes goodFFTSize gives) compiled for their
// size.  PAD: lines in the padded layout
// (lds_pad / lds_line) instead of n apart.
template <typename T, bool PAD = false>
__device__ __forceinline__ cplx<T>* fft_lds_c(cplx<T>* a, cplx<T>* b, int n, int nseg,
                                              const cplx<T>* tw, int twstride)
{
    switch (n) {
      case 64: return fft_fixed<T, 64, 1, PAD>(a, b, nseg, tw, twstride);
      case 96: return fft_fixed<T, 96, 1, PAD>(a, b, nseg, tw, twstride);
      case 128: return fft_fixed<T, 128, 1, PAD>(a, b, nseg, tw, twstride);
      case 192: return fft_fixed<T, 192, 1, PAD>(a, b, nseg, tw, twstride);
      case 256: return fft_fixed<T, 256, 1, PAD>(a, b, nseg, tw, twstride);
      case 384: return fft_fixed<T, 384, 1, PAD>(a, b, nseg, tw, twstride);
      case 512: return fft_fixed<T, 512, 1, PAD>(a, b, nseg, tw, twstride);
      case 768: return fft_fixed<T, 768, 1, PAD>(a, b, nseg, tw, twstride);
      case 1024: return fft_fixed<T, 1024, 1, PAD>(a, b, nseg, tw, twstride);
      case 1536: return fft_fixed<T, 1536, 1, PAD>(a, b, nseg, tw, twstride);
      case 2048: return fft_fixed<T, 2048, 1, PAD>(a, b, nseg, tw, twstride);
      case 3072: return fft_fixed<T, 3072, 1, PAD>(a, b, nseg, tw, twstride);
      default: return fft_lds<T, PAD>(a, b, n, nseg, tw, twstride);
    }
}

// Complex values between the columns of a stamp's column-transformed scratch (gsmi_stamp).
__device__ __forceinline__ int col_stride(const gsmi_stamp* st)
{
    return st->col_stride > st->ny ? st->col_stride : st->ny;
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs, each with its own L2.  This
// bijection of [0, G) gives each XCD one contiguous range of logical indices, so neighbouring
// tiles (adjacent columns, adjacent column chunks of one stamp) share an L2 and whole cache lines.
__device__ __forceinline__ int xcd_contiguous(int b, int G)
{
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, s = b >> 3;
    return (x < r) ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

__host__ __device__ inline int cols_per_group(int N)
{
    return N >= 1024 ? 1 : 1024 / N;
}
__host__ __device__ inline int rows_per_group(int N)
{
    int n2 = N / 2;
    return n2 >= 1024 ? 1 : 1024 / n2;
}

}  // namespace gsmi
