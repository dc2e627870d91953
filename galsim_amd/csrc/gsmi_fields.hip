// Parameter maps of the differentiable Exponential (x) Moffat model, fused into one kernel each
// way: parameter rows -> the lowered records' fields (written straight into the device records),
// and field gradients -> parameter gradients (the chain rule).  They restate on the device what
// the host lowering does with galsim/exponential.py (r0 = hlr / 1.6783469900166605),
// galsim/shear.py:365-386 (the shear matrix), galsim/moffat.py (rD from fwhm and beta) and the
// Moffat normalisation of src/SBMoffat.cpp:226-243; nothing here replaces a reference entry
// point.  One thread per stamp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/gsmi.h"
#include "launch.h"

namespace gsmi {
namespace {

constexpr int kF = GSMI_GRAD_FIELDS;
constexpr int kP = GSMI_GRAD_PER_LEAF;
constexpr int kUX = kP * GSMI_GRAD_MAXLEAF;
constexpr double kHlrToR0 = 1.6783469900166605;   // Exponential hlr / r0

// psi(x), x > 0: the recurrence psi(x) = psi(x + 1) - 1/x up to x >= 10, then the asymptotic
// series of Abramowitz & Stegun 6.3.18 (terms to x^-12; truncation < 1e-16 at x = 10).
__device__ double digamma(double x)
{
    double r = 0.;
    while (x < 10.) {
        r -= 1. / x;
        x += 1.;
    }
    const double i2 = 1. / (x * x);
    const double s = i2 * (1. / 12. - i2 * (1. / 120. - i2 * (1. / 252. - i2 * (1. / 240. -
                     i2 * (1. / 132. - i2 * (691. / 32760.))))));
    return r + log(x) - 0.5 / x - s;
}

// The Moffat record values that follow beta on the k^nu K_nu form (nu = beta - 1): the k = 0
// value 2^(nu-1) Gamma(nu) and the order-only constants of Temme's series for mu = nu - round(nu)
// (fact gam1, fact gam2, 1 / (2 Gamma(1+mu)), 1 / (2 Gamma(1-mu))), as moffat.temme_constants.
__device__ void moffat_aux(double beta, double* aux)
{
    const double nu = beta - 1.;
    aux[0] = exp((nu - 1.) * M_LN2 + lgamma(nu));
    const double mu = nu - floor(nu + 0.5);
    const double gampl = exp(-lgamma(1. + mu));
    const double gammi = exp(-lgamma(1. - mu));
    const bool small = fabs(mu) < 1e-4;
    const double m2 = mu * mu;
    const double gam1 = small ? -(0.5772156649015329 - 0.0420026350340952 * m2)
                              : (gammi - gampl) / (2. * mu);
    const double gam2 = small ? 1. - 0.6558780715202538 * m2 + 0.1665386113822915 * m2 * m2
                              : (gammi + gampl) / 2.;
    const double fact = small ? 1. + (M_PI * mu) * (M_PI * mu) / 6.
                              : M_PI * mu / sin(M_PI * mu);
    aux[1] = fact * gam1;
    aux[2] = fact * gam2;
    aux[3] = 0.5 / gampl;
    aux[4] = 0.5 / gammi;
}

// theta row: flux, hlr, g1, g2, dx, dy, psf_fwhm [, beta].  f: the stamp's fields (base where no
// parameter reaches).
__device__ void fields_of(const double* th, bool fit_beta, const double* base, double cx,
                          double cy, double inv_scale, double fwhm_to_rd, double* f)
{
    const double g1 = th[2], g2 = th[3];
    const double a = inv_scale / sqrt(1. - (g1 * g1 + g2 * g2));
#pragma unroll
    for (int k = 0; k < kF; ++k) f[k] = base[k];
    f[0] = th[0];
    f[1] = th[1] / kHlrToR0;
    f[2] = a * (1. + g1);
    f[3] = a * g2;
    f[4] = a * g2;
    f[5] = a * (1. - g1);
    if (fit_beta) {
        const double beta = th[7];
        f[kP] = 4. * exp(-lgamma(beta - 1.) - beta * M_LN2);
        f[kP + 1] = th[6] / (2. * sqrt(pow(2., 1. / beta) - 1.));
        f[kP + 6] = beta;
    } else {
        f[kP + 1] = th[6] * fwhm_to_rd;
    }
    f[kUX] = th[4] + cx;
    f[kUX + 1] = th[5] + cy;
}

__global__ void __launch_bounds__(256)
exp_moffat_fields_kernel(const double* __restrict__ theta, int n, int fit_beta,
                         const double* __restrict__ base, const double* __restrict__ cxcy,
                         double inv_scale, const double* __restrict__ fwhm_to_rd,
                         const int64_t* __restrict__ dst, const int64_t* __restrict__ aux_dst,
                         double* __restrict__ records)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ncol = fit_beta ? 8 : 7;
    const double* th = theta + size_t(i) * ncol;
    double f[kF];
    fields_of(th, fit_beta != 0, base + size_t(i) * kF, cxcy[2 * i], cxcy[2 * i + 1], inv_scale,
              fit_beta ? 0. : fwhm_to_rd[i], f);
    const int64_t* d = dst + size_t(i) * kF;
#pragma unroll
    for (int k = 0; k < kF; ++k)
        if (d[k] >= 0) records[d[k]] = f[k];
    if (fit_beta && aux_dst) {
        double aux[5];
        moffat_aux(th[7], aux);
#pragma unroll
        for (int k = 0; k < 5; ++k) records[aux_dst[size_t(i) * 5 + k]] = aux[k];
    }
}

// dL/dtheta from dL/dfields (rows in the plan's sorted order, inv[i] = row of stamp i).
__global__ void __launch_bounds__(256)
exp_moffat_grad_kernel(const double* __restrict__ theta, int n, int fit_beta,
                       const double* __restrict__ fwhm_to_rd, const double* __restrict__ gfields,
                       const int64_t* __restrict__ inv, double inv_scale,
                       double* __restrict__ dtheta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ncol = fit_beta ? 8 : 7;
    const double* th = theta + size_t(i) * ncol;
    const double* G = gfields + size_t(inv[i]) * kF;
    double* out = dtheta + size_t(i) * ncol;
    const double g1 = th[2], g2 = th[3];
    const double den2 = 1. - (g1 * g1 + g2 * g2);
    const double a = inv_scale / sqrt(den2);          // d a / d g = a g / den2
    const double b = a / den2;
    out[0] = G[0];
    out[1] = G[1] / kHlrToR0;
    // f2 = a (1 + g1), f3 = f4 = a g2, f5 = a (1 - g1)
    out[2] = G[2] * (a + b * (1. + g1) * g1) + (G[3] + G[4]) * (b * g2 * g1) +
             G[5] * (-a + b * (1. - g1) * g1);
    out[3] = G[2] * (b * (1. + g1) * g2) + (G[3] + G[4]) * (a + b * g2 * g2) +
             G[5] * (b * (1. - g1) * g2);
    out[4] = G[kUX];
    out[5] = G[kUX + 1];
    if (fit_beta) {
        const double beta = th[7];
        const double p = pow(2., 1. / beta);
        const double s = sqrt(p - 1.);
        out[6] = G[kP + 1] / (2. * s);
        const double mflux = 4. * exp(-lgamma(beta - 1.) - beta * M_LN2);
        // d mflux / d beta = -mflux (psi(beta - 1) + ln 2); d rD / d beta = fw p ln2 / (4 s^3 b^2)
        out[7] = G[kP] * (-mflux * (digamma(beta - 1.) + M_LN2)) +
                 G[kP + 1] * (th[6] * p * M_LN2 / (4. * s * s * s * beta * beta)) + G[kP + 6];
    } else {
        out[6] = G[kP + 1] * fwhm_to_rd[i];
    }
}

}  // namespace
}  // namespace gsmi

using namespace gsmi;

extern "C" {

int gsmi_exp_moffat_fields(const double* theta, int n, int fit_beta, const double* base,
                           const double* cxcy, double inv_scale, const double* fwhm_to_rd,
                           const int64_t* dst, const int64_t* aux_dst, double* records,
                           void* stream)
{
    if (n < 0 || !theta || !base || !cxcy || !dst || !records || (!fit_beta && !fwhm_to_rd)) {
        set_error("gsmi_exp_moffat_fields: bad arguments");
        return GSMI_E_ARG;
    }
    if (n == 0) return GSMI_OK;
    hipLaunchKernelGGL(exp_moffat_fields_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), theta, n, fit_beta, base, cxcy,
                       inv_scale, fwhm_to_rd, dst, aux_dst, records);
    return hip_check(hipGetLastError(), "exp_moffat_fields launch");
}

int gsmi_exp_moffat_grad(const double* theta, int n, int fit_beta, const double* fwhm_to_rd,
                         const double* gfields, const int64_t* inv, double inv_scale,
                         double* dtheta, void* stream)
{
    if (n < 0 || !theta || !gfields || !inv || !dtheta || (!fit_beta && !fwhm_to_rd)) {
        set_error("gsmi_exp_moffat_grad: bad arguments");
        return GSMI_E_ARG;
    }
    if (n == 0) return GSMI_OK;
    hipLaunchKernelGGL(exp_moffat_grad_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), theta, n, fit_beta, fwhm_to_rd,
                       gfields, inv, inv_scale, dtheta);
    return hip_check(hipGetLastError(), "exp_moffat_grad launch");
}

}  // extern "C"
