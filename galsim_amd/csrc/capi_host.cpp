// Host half of the C ABI: profile scalars and tables (see include/gsmi.h).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/gsmi.h"
#include "host_numerics.h"

namespace gsmi {
void set_error(const std::string& msg);
}

using namespace gsmi;

namespace {
template <class F>
int guarded(const char* name, F f)
{
    try {
        f();
        return GSMI_OK;
    } catch (const std::exception& e) {
        set_error(std::string(name) + ": " + e.what());
        std::string w = e.what();
        if (w.find("out of range") != std::string::npos) return GSMI_E_RANGE;
        return GSMI_E_SOLVE;
    }
}
bool is_finite(double x) { return std::isfinite(x); }
}  // namespace

extern "C" {

int gsmi_good_fft_size(int n) { return good_fft_size(n); }

int gsmi_sersic_info(double n, double trunc, const double* gsp, double* out, int* ntab)
{
    if (!out || !is_finite(n) || !(trunc >= 0.)) {
        set_error("gsmi_sersic_info: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_sersic_info", [&] {
        const SersicInfo& s = sersic_info(n, trunc, GSP::from(gsp));
        double v[11] = {s.maxk, s.stepk, s.hlr, s.b, s.ksq_min, s.ksq_max, s.highk_a,
                        s.highk_b, s.kderiv2, s.kderiv4, s.flux_fraction};
        for (int i = 0; i < 11; ++i) out[i] = v[i];
        if (ntab) *ntab = int(s.ft.args.size());
    });
}

int gsmi_sersic_table(double n, double trunc, const double* gsp, double* args, double* vals,
                      double* y2, int cap)
{
    if (!args || !vals || !y2 || !is_finite(n) || !(trunc >= 0.)) {
        set_error("gsmi_sersic_table: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_sersic_table", [&] {
        const SersicInfo& s = sersic_info(n, trunc, GSP::from(gsp));
        int m = int(s.ft.args.size());
        if (m > cap) throw std::runtime_error("table capacity too small");
        if (!s.ft.equal_spaced) throw std::runtime_error("Sersic table knots not equally spaced");
        for (int i = 0; i < m; ++i) {
            args[i] = s.ft.args[i];
            vals[i] = s.ft.vals[i];
            y2[i] = s.ft.y2[i];
        }
    });
}

int gsmi_sersic_kvalue(double n, double trunc, const double* gsp, const double* ksq, double* out,
                       int count)
{
    if (!ksq || !out || count < 0 || !(trunc >= 0.)) {
        set_error("gsmi_sersic_kvalue: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_sersic_kvalue", [&] {
        const SersicInfo& s = sersic_info(n, trunc, GSP::from(gsp));
        for (int i = 0; i < count; ++i) {
            double q = ksq[i];
            if (q >= s.ksq_max) out[i] = (s.highk_a + s.highk_b / std::sqrt(q)) / q;
            else if (q < s.ksq_min) out[i] = 1. + q * (s.kderiv2 + q * s.kderiv4);
            else out[i] = s.ft(0.5 * std::log(q)) / q;
        }
    });
}

int gsmi_sersic_hlr(double n, double flux_fraction, double* out)
{
    if (!out || !is_finite(n)) { set_error("gsmi_sersic_hlr: bad argument"); return GSMI_E_ARG; }
    return guarded("gsmi_sersic_hlr", [&] { *out = sersic_hlr(n, flux_fraction); });
}

int gsmi_sersic_integrated_flux(double n, double r, double* out)
{
    if (!out || !is_finite(n) || !(r >= 0.)) { set_error("bad argument"); return GSMI_E_ARG; }
    return guarded("gsmi_sersic_integrated_flux", [&] { *out = sersic_integrated_flux(n, r); });
}

int gsmi_sersic_truncated_scale(double n, double hlr, double trunc, double* out)
{
    if (!out || !is_finite(n) || !(hlr > 0.) || !(trunc > 0.)) {
        set_error("gsmi_sersic_truncated_scale: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_sersic_truncated_scale",
                   [&] { *out = sersic_truncated_scale(n, hlr, trunc); });
}

int gsmi_kolmogorov_info(const double* gsp, double* out)
{
    if (!out) { set_error("gsmi_kolmogorov_info: bad argument"); return GSMI_E_ARG; }
    return guarded("gsmi_kolmogorov_info", [&] {
        const KolmogorovInfo& k = kolmogorov_info(GSP::from(gsp));
        out[0] = k.maxk;
        out[1] = k.stepk;
    });
}

int gsmi_moffat_info(double beta, double scale_radius, double flux, const double* gsp, double* out)
{
    if (!out || !is_finite(beta) || !(scale_radius > 0)) {
        set_error("gsmi_moffat_info: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_moffat_info", [&] {
        MoffatInfo m = moffat_info(beta, scale_radius, flux, GSP::from(gsp));
        out[0] = m.maxk;
        out[1] = m.stepk;
        out[2] = m.knorm;
        out[3] = m.knorm2;
        out[4] = m.kind;
    });
}

int gsmi_exponential_cliprows(double r0, const double* gsp, const double* jac, int Nk, double dk,
                              int32_t* out)
{
    if (!out || Nk < 2 || Nk % 2 || !(r0 > 0) || !(dk > 0)) {
        set_error("gsmi_exponential_cliprows: bad argument");
        return GSMI_E_ARG;
    }
    GSP g = GSP::from(gsp);
    // SBExponentialImpl ctor (src/SBExponential.cpp:71-76)
    const double ksq_max = std::pow(g.kvalue_accuracy, -1. / 1.5) - 1.;
    const double kmax = std::sqrt(ksq_max);
    const int m = Nk / 2 + 1;
    const int half = Nk / 2;
    // SBProfile::drawK on bounds (0..Nk/2, -Nk/2..Nk/2) (src/SBProfile.cpp:383-425)
    const double x0 = 0. * dk;
    const double y0 = -half * dk;
    if (!jac || (jac[1] == 0. && jac[2] == 0.)) {
        // 1-d fill through FillQuadrant: quadrant rows ky = 0, dky, ... (src/SBProfile.cpp:447-506,
        // GetKValueRange1d 508-531)
        double A = jac ? jac[0] : 1., D = jac ? jac[3] : 1.;
        double dkx = (jac ? dk * A : dk) * r0;
        double dky = (jac ? dk * D : dk) * r0;
        double kx0 = (jac ? x0 * A : x0) * r0;
        double ky = 0. * r0;
        for (int q = 0; q <= half; ++q, ky += dky) {
            int i1, i2;
            if (std::abs(ky) >= kmax) {
                i1 = i2 = m;
            } else {
                double kysq = ky * ky;
                double dsq = ksq_max - kysq;
                if (kx0 * kx0 > dsq || (kx0 + m * dkx) * (kx0 + m * dkx) > dsq) {
                    double d = std::sqrt(dsq);
                    i1 = int(std::ceil((-kx0 - d) / dkx));
                    i2 = int(std::floor((-kx0 + d) / dkx));
                    if (i1 > i2) std::swap(i1, i2);
                    ++i2;
                    if (i2 <= 0 || i1 >= m) i1 = i2 = m;
                    else { if (i1 < 0) i1 = 0; if (i2 > m) i2 = m; }
                } else {
                    i1 = 0;
                    i2 = m;
                }
            }
            for (int j : {half - q, half + q}) {
                out[2 * j] = i1;
                out[2 * j + 1] = i2;
            }
        }
    } else {
        // general Jacobian: 2-d fill, GetKValueRange2d (src/SBProfile.cpp:533-588) including its
        // test of |ky0 + m + dky| (line 544)
        double A = jac[0], B = jac[1], C = jac[2], D = jac[3];
        double kx0 = (A * x0 + C * y0) * r0;
        double dkx = (A * dk) * r0;
        double dkxy = (C * dk) * r0;
        double ky0 = (B * x0 + D * y0) * r0;
        double dky = (D * dk) * r0;
        double dkyx = (B * dk) * r0;
        for (int j = 0; j <= Nk; ++j, kx0 += dkxy, ky0 += dky) {
            int i1, i2;
            if (std::abs(kx0) > kmax || std::abs(kx0 + m * dkx) > kmax ||
                std::abs(ky0) > kmax || std::abs(ky0 + m + dkyx) > kmax) {
                double a = dkx * dkx + dkyx * dkyx;
                double b = dkx * kx0 + dkyx * ky0;
                double c = kx0 * kx0 + ky0 * ky0 - ksq_max;
                double d = b * b - a * c;
                if (d <= 0.) {
                    i1 = i2 = m;
                } else {
                    d = std::sqrt(d);
                    i1 = int(std::ceil((-b - d) / a));
                    i2 = int(std::floor((-b + d) / a));
                    if (i1 > i2) std::swap(i1, i2);
                    ++i2;
                    if (i2 <= 0 || i1 >= m) i1 = i2 = m;
                    else { if (i1 < 0) i1 = 0; if (i2 > m) i2 = m; }
                }
            } else {
                i1 = 0;
                i2 = m;
            }
            out[2 * j] = i1;
            out[2 * j + 1] = i2;
        }
    }
    return GSMI_OK;
}

int gsmi_moffat_scale_radius_from_hlr(double hlr, double beta, double* out)
{
    if (!out || !(hlr > 0) || !is_finite(beta)) { set_error("bad argument"); return GSMI_E_ARG; }
    *out = moffat_scale_radius_from_hlr(hlr, beta);
    return GSMI_OK;
}

int gsmi_moffat_scale_radius_from_hlr_trunc(double hlr, double trunc, double beta, double* out)
{
    if (!out || !(hlr > 0) || !(trunc > 0) || !is_finite(beta)) {
        set_error("gsmi_moffat_scale_radius_from_hlr_trunc: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_moffat_scale_radius_from_hlr_trunc",
                   [&] { *out = moffat_scale_radius_from_hlr_trunc(hlr, trunc, beta); });
}

int gsmi_moffat_trunc_info(double beta, double scale_radius, double trunc, const double* gsp,
                           double* out)
{
    if (!out || !is_finite(beta) || !(scale_radius > 0) || !(trunc > 0)) {
        set_error("gsmi_moffat_trunc_info: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_moffat_trunc_info", [&] {
        const MoffatTruncInfo& m = moffat_trunc_info(beta, trunc / scale_radius, GSP::from(gsp));
        out[0] = m.maxk_rD / scale_radius;
        out[1] = m.stepk_rD / scale_radius;
        out[2] = m.flux_factor;
        out[3] = m.hlr_rD * scale_radius;
        out[4] = double(m.ft.args.size());
        out[5] = m.dk;
    });
}

int gsmi_moffat_trunc_table(double beta, double scale_radius, double trunc, const double* gsp,
                            double* args, double* vals, double* y2, int cap)
{
    if (!args || !vals || !y2 || !is_finite(beta) || !(scale_radius > 0) || !(trunc > 0)) {
        set_error("gsmi_moffat_trunc_table: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_moffat_trunc_table", [&] {
        const MoffatTruncInfo& m = moffat_trunc_info(beta, trunc / scale_radius, GSP::from(gsp));
        if (int(m.ft.args.size()) > cap) throw std::runtime_error("table capacity too small");
        std::copy(m.ft.args.begin(), m.ft.args.end(), args);
        std::copy(m.ft.vals.begin(), m.ft.vals.end(), vals);
        std::copy(m.ft.y2.begin(), m.ft.y2.end(), y2);
    });
}

}
