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

}  // namespace gsmi
