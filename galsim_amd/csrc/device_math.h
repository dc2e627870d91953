// Device-side math for the galsim_amd kernels: complex helpers, the modified Bessel function
// K_nu (Temme's series for x <= 2, Steed's continued fraction CF2 for x > 2, then upward
// recurrence in the order), and the leaf k-space formulas.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gsmi.h"

namespace gsmi {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }

__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dexp(float x) { return __expf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dlog(float x) { return __logf(x); }
__device__ __forceinline__ void dsincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void dsincos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ double dsinpi(double x) { return sinpi(x); }
__device__ __forceinline__ float dsinpi(float x) { return sinpif(x); }

// Exact integer quotient t / d for 0 <= t, 1 <= d, t + d < 2^21, from inv_d = 1.f / d: the
// product (t + 1/2) / d lies at least 1/(2d) away from an integer and its float error is below
// 2^-23 (q + 1), so truncation gives q.  The kernels' index arithmetic divides by FFT sizes and
// tile widths that are not powers of two (3 * 2^k); this replaces ~40-instruction integer
// divisions by one multiply.
__device__ __forceinline__ int fdiv(int t, float inv_d)
{
    return int((float(t) + 0.5f) * inv_d);
}

// Division for the leaf formulas: exact in FP64 (the FP64 path must reproduce the reference),
// the hardware reciprocal (1 ulp) in FP32.
__device__ __forceinline__ double ddiv(double a, double b) { return a / b; }
__device__ __forceinline__ float ddiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// 1/Gamma(1+mu) +- 1/Gamma(1-mu) combinations for Temme's series (|mu| <= 1/2).  For small mu
// the Taylor coefficients of 1/Gamma(1+z) (Abramowitz & Stegun 6.1.34) avoid cancellation.
__device__ inline void temme_gammas(double mu, double* gam1, double* gam2, double* gampl,
                                    double* gammi)
{
    *gampl = 1. / tgamma(1. + mu);
    *gammi = 1. / tgamma(1. - mu);
    if (fabs(mu) < 1e-4) {
        const double c2 = 0.5772156649015329, c3 = -0.6558780715202538,
                     c4 = -0.0420026350340952, c5 = 0.1665386113822915;
        double m2 = mu * mu;
        *gam1 = -(c2 + c4 * m2);
        *gam2 = 1. + c3 * m2 + c5 * m2 * m2;
    } else {
        *gam1 = (*gammi - *gampl) / (2. * mu);
        *gam2 = (*gammi + *gampl) / 2.;
    }
}

// K_nu(x) for nu >= 0, x > 0 (Temme 1975; Steed's algorithm for CF2).  With km1 != nullptr
// also returns K_{nu-1}(x).
__device__ inline double bessel_k(double nu, double x, double* km1 = nullptr)
{
    const double EPS = 1e-16;
    const int MAXIT = 10000;
    int nl = int(nu + 0.5);
    double mu = nu - nl;              // |mu| <= 1/2
    double mu2 = mu * mu;
    double xi = 1. / x, xi2 = 2. * xi;
    double rkmu, rk1;
    if (x < 2.) {
        double x2 = 0.5 * x;
        double pimu = M_PI * mu;
        double fact = (fabs(pimu) < 1e-16) ? 1. : pimu / sin(pimu);
        double d = -log(x2);
        double e = mu * d;
        double fact2 = (fabs(e) < 1e-16) ? 1. : sinh(e) / e;
        double gam1, gam2, gampl, gammi;
        temme_gammas(mu, &gam1, &gam2, &gampl, &gammi);
        double ff = fact * (gam1 * cosh(e) + gam2 * fact2 * d);
        double sum = ff;
        e = exp(e);
        double p = 0.5 * e / gampl;
        double q = 0.5 / (e * gammi);
        double c = 1.;
        d = x2 * x2;
        double sum1 = p;
        for (int i = 1; i <= MAXIT; ++i) {
            ff = (i * ff + p + q) / (i * i - mu2);
            c *= d / i;
            p /= (i - mu);
            q /= (i + mu);
            double del = c * ff;
            sum += del;
            double del1 = c * (p - i * ff);
            sum1 += del1;
            if (fabs(del) < fabs(sum) * EPS) break;
        }
        rkmu = sum;
        rk1 = sum1 * xi2;
    } else {
        double b = 2. * (1. + x);
        double d = 1. / b;
        double h = d, delh = d;
        double q1 = 0., q2 = 1.;
        double a1 = 0.25 - mu2;
        double q = a1, c = a1;
        double a = -a1;
        double s = 1. + q * delh;
        for (int i = 1; i <= MAXIT; ++i) {
            a -= 2 * i;
            c = -a * c / (i + 1.);
            double qnew = (q1 - b * q2) / a;
            q1 = q2;
            q2 = qnew;
            q += c * qnew;
            b += 2.;
            d = 1. / (b + a * d);
            delh = (b * d - 1.) * delh;
            h += delh;
            double dels = q * delh;
            s += dels;
            if (fabs(dels / s) < EPS) break;
        }
        rkmu = sqrt(M_PI / (2. * x)) * exp(-x) / s;
        rk1 = rkmu * (mu + x + 0.5 - a1 * h) * xi;
    }
    // upward recurrence K_{m+1} = K_{m-1} + (2m/x) K_m from m = mu+1
    double kmu = rkmu, kmu1 = rk1, kprev = 0.;
    for (int i = 1; i <= nl; ++i) {
        double knext = (mu + i) * xi2 * kmu1 + kmu;
        kprev = kmu;
        kmu = kmu1;
        kmu1 = knext;
    }
    if (km1) {
        // nu < 1/2: K_{nu-1} = K_{nu+1} - (2 nu / x) K_nu (loses about log10(1/x^(2 nu)) digits
        // as x -> 0, a digit or two over the k range of a drawn stamp)
        *km1 = (nl >= 1) ? kprev : kmu1 - mu * xi2 * kmu;
    }
    return kmu;
}

// K_nu(x) in FP32 for nu = nl + mu (nl >= 0 integer, |mu| <= 1/2), x > 0, by the same method as
// bessel_k (Temme's series for x < 2, Steed's CF2 above, upward recurrence in the order).  The
// mu-only constants of Temme's series come in precomputed (in FP64, once per profile):
// c0 = fact gam1, c1 = fact gam2, c2 = Gamma(1 + mu) / 2, c3 = Gamma(1 - mu) / 2 with
// fact = pi mu / sin(pi mu).  With km1 != nullptr also returns K_{nu-1}(x).
__device__ inline float bessel_k_f32(int nl, float mu, float c0, float c1, float c2, float c3,
                                     float x, float* km1)
{
    const float EPS = 1e-7f;
    const int MAXIT = 1000;
    const float mu2 = mu * mu;
    const float xi = 1.f / x, xi2 = 2.f * xi;
    float rkmu, rk1;
    if (x < 2.f) {
        const float x2 = 0.5f * x;
        float d = -__logf(x2);
        float e = mu * d;
        const float fact2 = (fabsf(e) < 1e-4f) ? 1.f + e * e * (1.f / 6.f) : sinhf(e) / e;
        float ff = c0 * coshf(e) + c1 * fact2 * d;
        float sum = ff;
        e = __expf(e);
        float p = c2 * e;
        float q = c3 / e;
        float c = 1.f;
        d = x2 * x2;
        float sum1 = p;
        for (int i = 1; i <= MAXIT; ++i) {
            const float fi = float(i);
            ff = (fi * ff + p + q) / (fi * fi - mu2);
            c *= d / fi;
            p /= (fi - mu);
            q /= (fi + mu);
            const float del = c * ff;
            sum += del;
            sum1 += c * (p - fi * ff);
            if (fabsf(del) < fabsf(sum) * EPS) break;
        }
        rkmu = sum;
        rk1 = sum1 * xi2;
    } else {
        float b = 2.f * (1.f + x);
        float d = 1.f / b;
        float h = d, delh = d;
        float q1 = 0.f, q2 = 1.f;
        const float a1 = 0.25f - mu2;
        float q = a1, c = a1;
        float a = -a1;
        float s = 1.f + q * delh;
        for (int i = 1; i <= MAXIT; ++i) {
            a -= 2 * i;
            c = -a * c / (i + 1.f);
            const float qnew = (q1 - b * q2) / a;
            q1 = q2;
            q2 = qnew;
            q += c * qnew;
            b += 2.f;
            d = 1.f / (b + a * d);
            delh = (b * d - 1.f) * delh;
            h += delh;
            const float dels = q * delh;
            s += dels;
            if (fabsf(dels / s) < EPS) break;
        }
        rkmu = sqrtf(float(M_PI) / (2.f * x)) * __expf(-x) / s;
        rk1 = rkmu * (mu + x + 0.5f - a1 * h) * xi;
    }
    float kmu = rkmu, kmu1 = rk1, kprev = 0.f;
    for (int i = 1; i <= nl; ++i) {
        const float knext = (mu + i) * xi2 * kmu1 + kmu;
        kprev = kmu;
        kmu = kmu1;
        kmu1 = knext;
    }
    if (km1) *km1 = (nl >= 1) ? kprev : kmu1 - mu * xi2 * kmu;
    return kmu;
}

// K0(x) and K1(x) in FP32 for x > 0 from the polynomial approximations of Abramowitz & Stegun,
// Handbook of Mathematical Functions, 9.8.1, 9.8.3, 9.8.5-9.8.8 (absolute error below 2.2e-7 in
// x^(1/2) e^x K(x) for x >= 2, below 1e-8 in K0 and x K1 for x <= 2).  Returns K0 and x K1, the
// form the Moffat kernels need (x K1 -> 1 as x -> 0).
__device__ inline void bessel_k0_xk1_f32(float x, float* k0, float* xk1)
{
    if (x <= 2.f) {
        const float t = x * (1.f / 3.75f);
        const float t2 = t * t;
        const float i0 = 1.f + t2 * (3.5156229f + t2 * (3.0899424f + t2 * (1.2067492f +
                         t2 * (0.2659732f + t2 * (0.0360768f + t2 * 0.0045813f)))));
        const float i1_x = 0.5f + t2 * (0.87890594f + t2 * (0.51498869f + t2 * (0.15084934f +
                           t2 * (0.02658733f + t2 * (0.00301532f + t2 * 0.00032411f)))));
        const float y = 0.25f * x * x;
        const float lg = __logf(0.5f * x);
        *k0 = -lg * i0 + (-0.57721566f + y * (0.42278420f + y * (0.23069756f + y * (0.03488590f +
              y * (0.00262698f + y * (0.00010750f + y * 0.0000074f))))));
        *xk1 = x * x * lg * i1_x + (1.f + y * (0.15443144f + y * (-0.67278579f +
               y * (-0.18156897f + y * (-0.01919402f + y * (-0.00110404f + y * -0.00004686f))))));
    } else {
        const float z = 2.f / x;
        const float e = __expf(-x) * rsqrtf(x);
        *k0 = e * (1.25331414f + z * (-0.07832358f + z * (0.02189568f + z * (-0.01062446f +
              z * (0.00587872f + z * (-0.00251540f + z * 0.00053208f))))));
        *xk1 = x * e * (1.25331414f + z * (0.23498619f + z * (-0.03655620f + z * (0.01504268f +
               z * (-0.00780353f + z * (0.00325614f + z * -0.00068245f))))));
    }
}

// Moffat k-profiles with integer Bessel order in FP32: g(k) = k K1 (beta = 2), k^2 K2 (3),
// k^3 K3 (4), from K0 and k K1 by the recurrence K_{n+1} = K_{n-1} + (2n/k) K_n; dg = g'(k)/(2k)
// = -K0/2, -k K1/2, -k^2 K2/2.  k = 0 gives g(0) = 1, 2, 8 and dg = 0.
__device__ inline void moffat_int_kind_f32(int kind, float k, float* g, float* dg)
{
    if (k == 0.f) {
        *g = (kind == 2) ? 1.f : (kind == 3) ? 2.f : 8.f;
        *dg = 0.f;
        return;
    }
    float k0, xk1;
    bessel_k0_xk1_f32(k, &k0, &xk1);
    const float k2k2 = k * k * k0 + 2.f * xk1;          // k^2 K2
    if (kind == 2) {
        *g = xk1;
        *dg = -0.5f * k0;
    } else if (kind == 3) {
        *g = k2k2;
        *dg = -0.5f * xk1;
    } else {
        *g = k * k * xk1 + 4.f * k2k2;                   // k^3 K3
        *dg = -0.5f * k2k2;
    }
}

}  // namespace gsmi
