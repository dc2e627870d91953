// Host-side profile numerics for galsim_amd: the FP64 scalars and lookup tables that size a
// stamp and feed the device kernels (stepK/maxK, Sersic Hankel tables, Kolmogorov radial table,
// Moffat maxK), plus the root finder, spline and Hankel machinery they are built from.
#pragma once

#include <cstdint>
#include <vector>

namespace gsmi {

// GSParams as a flat record, in the order of galsim/gsparams.py:126-131.
struct GSP {
    double minimum_fft_size, maximum_fft_size, folding_threshold, stepk_minimum_hlr,
        maxk_threshold, kvalue_accuracy, xvalue_accuracy, table_spacing, realspace_relerr,
        realspace_abserr, integration_relerr, integration_abserr, shoot_accuracy;
    static GSP from(const double* g);
    bool operator<(const GSP& o) const;
};

double expd(double x);
int good_fft_size(int n);
double gamma_p(double a, double x);

// Natural cubic spline on (args, vals).
struct Spline {
    std::vector<double> args, vals, y2;
    double da = 0.;
    bool equal_spaced = true;
    void build();
    double lookup(double a) const;
    double operator()(double a) const;   // 0 outside [args.front(), args.back()]
};

// Order-0 Hankel transform int_0^inf f(r) J0(k r) r dr by the Ogata (2005) quadrature.
typedef double (*RadialFn)(double r, const void* ctx);
double hankel_inf(RadialFn f, const void* ctx, double k, double relerr, double abserr);

// SersicInfo in units of r0; trunc (in units of r0) > 0 for a truncated profile.
struct SersicInfo {
    double n = 0., trunc = 0., invn = 0., gamma2n = 0., flux_fraction = 1.;
    double maxk = 0., stepk = 0., hlr = 0., b = 0.;
    double kderiv2 = 0., kderiv4 = 0., ksq_min = 0., ksq_max = 0., highk_a = 0., highk_b = 0.;
    Spline ft;    // f(k) k^2 as a function of log k
};
const SersicInfo& sersic_info(double n, double trunc, const GSP& gsp);
double sersic_hlr(double n, double flux_fraction);
double sersic_integrated_flux(double n, double r);
double sersic_truncated_scale(double n, double hlr, double trunc);

struct KolmogorovInfo {
    double maxk = 0., stepk = 0.;
    std::vector<double> xargs, xvals, xy2;   // the x-space radial table (natural cubic spline)
};
const KolmogorovInfo& kolmogorov_info(const GSP& gsp);

// Moffat (untruncated): kind code of the k-space formula, normalisations and sizes.
struct MoffatInfo {
    int kind = 0;           // 15, 2, 25, 3, 35, 4 or 0 (generic beta)
    double knorm = 0., knorm2 = 0., maxk = 0., stepk = 0., rD = 0.;
};
MoffatInfo moffat_info(double beta, double rD, double flux, const GSP& gsp);
double moffat_kv(int kind, double beta, double knorm2, double flux_over_knorm, double ksq);
double moffat_scale_radius_from_hlr(double re, double beta);

// Truncated Moffat (trunc > 0): the Hankel-transform table kV_trunc on x = (k rD)^2 and the sizes,
// in units of rD (maxk * rD, stepk * rD); depends on (beta, trunc / rD, gsparams), cached.
struct MoffatTruncInfo {
    double flux_factor = 0., maxRrD = 0., maxk_rD = 0., stepk_rD = 0., hlr_rD = 0., dk = 0.;
    Spline ft;
};
const MoffatTruncInfo& moffat_trunc_info(double beta, double maxRrD, const GSP& gsp);
double moffat_scale_radius_from_hlr_trunc(double re, double rm, double beta);

double bessel_k(double nu, double x);

// Spergel (src/SBSpergel.cpp:529-653): sizes in units of the scale radius r0, and the clipping
// radius of the k-space fill, ksq_max = kvalue_accuracy^(-1/(nu+1)) - 1 (in (k r0)^2); cached
// per (nu, gsparams).
struct SpergelInfo {
    double maxk = 0., stepk = 0., hlr = 0., ksq_max = 0.;
};
const SpergelInfo& spergel_info(double nu, const GSP& gsp);
// Radius (in r0) enclosing the fraction f of the flux: Brent on the closed-form enclosed flux.
double spergel_flux_radius(double nu, double f);
// Enclosed flux fraction within u r0: 1 - 2 (1+nu) (u/2)^(nu+1) K_{nu+1}(u) / Gamma(nu+2).
double spergel_integrated_flux(double nu, double u);

// GetKValueRange1d / 2d (src/SBProfile.cpp:508-588) of a radial profile clipped at
// (k r0)^2 > ksq_max, on the stored rows of an (Nk/2+1) x (Nk+1) k-image with spacing dk: the
// [i1, i2) range of each row, out[2 j], out[2 j + 1] for j = 0..Nk.
void kclip_rows(double ksq_max, double r0, const double* jac, int Nk, double dk, int32_t* out);

// InclinedExponential (src/SBInclinedExponential.cpp): sizes in units of the scale radius and the
// k-value thresholds, for |cos i| and hs = pi/2 h0 |sin i| / r0.  The maximum k and the clipping
// radius come from the reference's solves (Brent, then bisection) on the k-value along the axes.
struct InclinedExpInfo {
    double maxk = 0., stepk = 0., ksq_min = 0., ksq_max = 0.;
};
InclinedExpInfo inclined_exp_info(double cosi, double hs, const GSP& gsp);
double inclined_exp_kvalue(double kx, double ky, double cosi, double hs, double ksq_min,
                           double ksq_max);

// InclinedSersic (src/SBInclinedSersic.cpp): the Sersic k-value (SersicInfo::kValue, k in 1/r0)
// at (kx, ky cos i) times the same vertical factor; maxk (in 1/r0) and the clipping radius from
// the reference's solves, ksq_min of the vertical factor's Taylor branch.
struct InclinedSersicInfo {
    double maxk = 0., ksq_min = 0., ksq_max = 0.;
};
double sersic_kvalue(const SersicInfo& s, double ksq);
InclinedSersicInfo inclined_sersic_info(double n, double trunc, double cosi, double hs,
                                        const GSP& gsp);

// Sine integral Si(x) = int_0^x sin(t)/t dt: its power series for |x| <= 4, else
// pi/2 + Im E1(i x) with E1's continued fraction (Abramowitz & Stegun 5.1.22, 5.2.21-23)
// evaluated by the modified Lentz method.
double sine_integral(double x);

// Lanczos(n, conserve_dc) (src/Interpolant.cpp:719-899): K(1..5), the transform's cut umax and
// its spline table on u = 0, du, 2 du, ... as the reference builds it.
struct LanczosInfo {
    double K[6] = {0., 0., 0., 0., 0., 0.};
    double umax = 0.;
    Spline utab;
};
const LanczosInfo& lanczos_info(int n, bool conserve_dc, const GSP& gsp);

}  // namespace gsmi
