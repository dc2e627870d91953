// Thin extern "C" driver over the reference's own profile numerics, compiled from the reference
// sources by oracle/Makefile into oracle/_ref/libgsref.so.
//
// TEST INFRASTRUCTURE ONLY: it exposes the reference's SersicInfo (src/SBSersic.cpp:223-537),
// KolmogorovInfo (src/SBKolmogorov.cpp:337-419), the Sersic/Moffat radius helpers, the Hankel
// transforms and the special functions (src/math/*.cpp) so the oracle's restatement can be
// pinned against them.
// The reference's image code (src/Image.cpp) needs Eigen and FFTW, which this image lacks, so the
// SBProfile draw path itself is not linked; --gc-sections drops the unreferenced code that would
// need it.
#include <functional>
#include <cmath>
#include "GSParams.h"
#include "SBSersic.h"
#include "SBSersicImpl.h"
#include "SBKolmogorovImpl.h"
#include "SBMoffat.h"
#include "SBProfileImpl.h"
#include "math/Bessel.h"
#include "math/Gamma.h"
#include "math/Hankel.h"
#include "fmath/fmath.hpp"
#include "Interpolant.h"
#include "math/Sinc.h"

using namespace galsim;

namespace {
GSParamsPtr make_gsp(const double* g)
{
    if (!g) return GSParamsPtr(GSParams());
    return GSParamsPtr(GSParams(int(g[0]), int(g[1]), g[2], g[3], g[4], g[5], g[6], g[7], g[8],
                                g[9], g[10], g[11], g[12]));
}
}

extern "C" {

// out[0]=maxK, out[1]=stepK, out[2]=HLR (all in units of r0), out[3]=XNorm
int gsref_sersic_info(double n, double trunc, const double* gsp, double* out)
{
    try {
        SersicInfo info(n, trunc, make_gsp(gsp));
        out[0] = info.maxK();
        out[1] = info.stepK();
        out[2] = info.getHLR();
        out[3] = info.getXNorm();
        return 0;
    } catch (...) { return -1; }
}

int gsref_sersic_kvalue(double n, double trunc, const double* gsp, const double* ksq,
                        double* out, int count)
{
    try {
        SersicInfo info(n, trunc, make_gsp(gsp));
        for (int i = 0; i < count; ++i) out[i] = info.kValue(ksq[i]);
        return 0;
    } catch (...) { return -1; }
}

double gsref_sersic_hlr(double n, double flux_fraction) { return SersicHLR(n, flux_fraction); }
double gsref_sersic_integrated_flux(double n, double r) { return SersicIntegratedFlux(n, r); }
double gsref_sersic_truncated_scale(double n, double hlr, double trunc)
{ return SersicTruncatedScale(n, hlr, trunc); }

// out[0]=maxK, out[1]=stepK in units of k0
int gsref_kolmogorov_info(const double* gsp, double* out)
{
    try {
        KolmogorovInfo info(make_gsp(gsp));
        out[0] = info.maxK();
        out[1] = info.stepK();
        return 0;
    } catch (...) { return -1; }
}

double gsref_moffat_scale_radius_from_hlr(double re, double rm, double beta)
{ return MoffatCalculateScaleRadiusFromHLR(re, rm, beta); }

// The reference's hankel_trunc (src/math/Hankel.cpp: GKP integration over [0, rmax] split at
// the first zeros of J0) of the truncated Moffat's radial profile (1 + r^2)^-beta, with the
// reference's generic power fmath::expd(-beta log x) (src/SBMoffat.cpp:46-47, 272): the integral
// that setupFT (src/SBMoffat.cpp:565-626) tabulates.
double gsref_hankel_moffat_trunc(double beta, double k, double rmax, double relerr,
                                 double abserr)
{
    std::function<double(double)> f = [beta](double r) {
        return fmath::expd(-beta * std::log(1. + r * r));
    };
    return math::hankel_trunc(f, k, 0., rmax, relerr, abserr);
}

double gsref_bessel_k(double nu, double x) { return math::cyl_bessel_k(nu, x); }
double gsref_bessel_j(double nu, double x) { return math::cyl_bessel_j(nu, x); }
double gsref_bessel_y(double nu, double x) { return math::cyl_bessel_y(nu, x); }
double gsref_bessel_root(double nu, int s) { return math::getBesselRoot(nu, s); }
double gsref_gamma_p(double a, double x) { return math::gamma_p(a, x); }

// Hankel transform of exp(-r^(1/n)) (the Sersic radial function, SBSersic.cpp:378-385)
double gsref_hankel_sersic(double invn, double k, double relerr, double abserr)
{
    std::function<double(double)> f = [invn](double r) { return std::exp(-std::pow(r, invn)); };
    return math::hankel_inf(f, k, 0., relerr, abserr);
}

// Hankel transform of exp(-k^(5/3)) (KolmKValue, SBKolmogorov.cpp:262-266) at radius r
double gsref_hankel_kolmogorov(double r, double relerr, double abserr)
{
    std::function<double(double)> f = [](double k) {
        return fmath::expd(-fmath::expd(5./3. * std::log(k))); };
    return math::hankel_inf(f, r, 0., relerr, abserr);
}

double gsref_expd(double x) { return fmath::expd(x); }

// GetKValueRange1d / GetKValueRange2d (src/SBProfile.cpp:508-588): the row range [i1, i2) of a
// k-image row that the clipped fills (SBExponential, SBSpergel) evaluate; out = {i1, i2}, and
// for the 1-d form out_kysq = ky^2 when the reference sets it.
void gsref_kvalue_range_1d(int m, double kmax, double ksqmax, double kx0, double dkx, double ky,
                           int* out, double* out_kysq)
{
    double kysq = -1.;
    GetKValueRange1d(out[0], out[1], m, kmax, ksqmax, kx0, dkx, ky, kysq);
    *out_kysq = kysq;
}

void gsref_kvalue_range_2d(int m, double kmax, double ksqmax, double kx0, double dkx, double ky0,
                           double dky, int* out)
{
    GetKValueRange2d(out[0], out[1], m, kmax, ksqmax, kx0, dkx, ky0, dky);
}


// math::Si (src/math/Sinc.cpp:101-170)
double gsref_si(double x) { return math::Si(x); }

// Interpolant xval / uval / ranges (src/Interpolant.cpp): code 0 Nearest, 1 Linear, 2 Cubic,
// 3 Quintic, 4 Lanczos(n, conserve_dc); out = {xrange, urange, ixrange}
namespace {
std::unique_ptr<Interpolant> make_interp(int code, int n, int dc, const double* gsp)
{
    GSParams g = *make_gsp(gsp);
    switch (code) {
      case 0: return std::unique_ptr<Interpolant>(new Nearest(g));
      case 1: return std::unique_ptr<Interpolant>(new Linear(g));
      case 2: return std::unique_ptr<Interpolant>(new Cubic(g));
      case 3: return std::unique_ptr<Interpolant>(new Quintic(g));
      default: return std::unique_ptr<Interpolant>(new Lanczos(n, dc != 0, g));
    }
}
}

int gsref_interp_eval(int code, int n, int dc, const double* gsp, const double* x, int count,
                      double* xval, double* uval, double* ranges)
{
    try {
        std::unique_ptr<Interpolant> I = make_interp(code, n, dc, gsp);
        for (int i = 0; i < count; ++i) {
            xval[i] = I->xval(x[i]);
            uval[i] = I->uval(x[i]);
        }
        ranges[0] = I->xrange();
        ranges[1] = I->urange();
        ranges[2] = I->ixrange();
        return 0;
    } catch (...) {
        return 1;
    }
}

}
