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

int gsmi_kolmogorov_xtable(const double* gsp, double* args, double* vals, double* y2, int n)
{
    int count = -1;
    int rc = guarded("gsmi_kolmogorov_xtable", [&] {
        const KolmogorovInfo& k = kolmogorov_info(GSP::from(gsp));
        count = int(k.xargs.size());
        if (args && vals && y2 && n >= count)
            for (int i = 0; i < count; ++i) {
                args[i] = k.xargs[i];
                vals[i] = k.xvals[i];
                y2[i] = k.xy2[i];
            }
    });
    return rc ? rc : count;
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
    kclip_rows(ksq_max, r0, jac, Nk, dk, out);
    return GSMI_OK;
}

int gsmi_kclip_rows(double ksq_max, double r0, const double* jac, int Nk, double dk,
                    int32_t* out)
{
    if (!out || Nk < 2 || Nk % 2 || !(r0 > 0) || !(dk > 0) || !(ksq_max > 0)) {
        set_error("gsmi_kclip_rows: bad argument");
        return GSMI_E_ARG;
    }
    kclip_rows(ksq_max, r0, jac, Nk, dk, out);
    return GSMI_OK;
}

int gsmi_spergel_info(double nu, const double* gsp, double* out)
{
    if (!out || !is_finite(nu)) {
        set_error("gsmi_spergel_info: bad argument");
        return GSMI_E_ARG;
    }
    if (!(nu >= -0.85 && nu <= 4.0)) {
        set_error("gsmi_spergel_info: Spergel index out of range [-0.85, 4]");
        return GSMI_E_RANGE;
    }
    return guarded("gsmi_spergel_info", [&] {
        const SpergelInfo& s = spergel_info(nu, GSP::from(gsp));
        out[0] = s.maxk;
        out[1] = s.stepk;
        out[2] = s.hlr;
        out[3] = s.ksq_max;
    });
}

int gsmi_inclined_exp_info(double cosi, double hs, const double* gsp, double* out)
{
    if (!out || !(cosi >= 0. && cosi <= 1.) || !(hs >= 0.) || !is_finite(hs)) {
        set_error("gsmi_inclined_exp_info: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_inclined_exp_info", [&] {
        InclinedExpInfo s = inclined_exp_info(cosi, hs, GSP::from(gsp));
        out[0] = s.maxk;
        out[1] = s.stepk;
        out[2] = s.ksq_min;
        out[3] = s.ksq_max;
    });
}

int gsmi_inclined_exp_kvalue(double cosi, double hs, double ksq_min, double ksq_max,
                             const double* kx, const double* ky, double* out, int count)
{
    if (!kx || !ky || !out || count < 0) {
        set_error("gsmi_inclined_exp_kvalue: bad argument");
        return GSMI_E_ARG;
    }
    for (int i = 0; i < count; ++i)
        out[i] = inclined_exp_kvalue(kx[i], ky[i], cosi, hs, ksq_min, ksq_max);
    return GSMI_OK;
}

int gsmi_inclined_sersic_info(double n, double trunc, double cosi, double hs, const double* gsp,
                              double* out)
{
    if (!out || !is_finite(n) || !(trunc >= 0.) || !(cosi >= 0. && cosi <= 1.) || !(hs >= 0.) ||
        !is_finite(hs)) {
        set_error("gsmi_inclined_sersic_info: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_inclined_sersic_info", [&] {
        InclinedSersicInfo s = inclined_sersic_info(n, trunc, cosi, hs, GSP::from(gsp));
        out[0] = s.maxk;
        out[1] = s.ksq_min;
        out[2] = s.ksq_max;
    });
}

int gsmi_spergel_flux_radius(double nu, double f, double* out)
{
    if (!out || !is_finite(nu) || !(f > 0. && f < 1.)) {
        set_error("gsmi_spergel_flux_radius: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_spergel_flux_radius", [&] { *out = spergel_flux_radius(nu, f); });
}

int gsmi_spergel_integrated_flux(double nu, double u, double* out)
{
    if (!out || !is_finite(nu) || !(u >= 0.)) {
        set_error("gsmi_spergel_integrated_flux: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_spergel_integrated_flux",
                   [&] { *out = spergel_integrated_flux(nu, u); });
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


int gsmi_sine_integral(const double* x, double* out, int n)
{
    if ((!x || !out) && n > 0) {
        set_error("gsmi_sine_integral: bad argument");
        return GSMI_E_ARG;
    }
    for (int i = 0; i < n; ++i) out[i] = sine_integral(x[i]);
    return GSMI_OK;
}

int gsmi_lanczos_info(int n, int conserve_dc, const double* gsp, double* out, int* ntab)
{
    if (!out || n < 1) {
        set_error("gsmi_lanczos_info: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_lanczos_info", [&] {
        const LanczosInfo& L = lanczos_info(n, conserve_dc != 0, GSP::from(gsp));
        for (int j = 0; j < 5; ++j) out[j] = L.K[j + 1];
        out[5] = L.umax;
        if (ntab) *ntab = int(L.utab.args.size());
    });
}

int gsmi_lanczos_utable(int n, int conserve_dc, const double* gsp, double* args, double* vals,
                        double* y2, int cap)
{
    if (!args || !vals || !y2 || n < 1) {
        set_error("gsmi_lanczos_utable: bad argument");
        return GSMI_E_ARG;
    }
    return guarded("gsmi_lanczos_utable", [&] {
        const LanczosInfo& L = lanczos_info(n, conserve_dc != 0, GSP::from(gsp));
        const int m = int(L.utab.args.size());
        if (cap < m) throw std::runtime_error("table out of range of the buffer");
        for (int i = 0; i < m; ++i) {
            args[i] = L.utab.args[i];
            vals[i] = L.utab.vals[i];
            y2[i] = L.utab.y2[i];
        }
    });
}

}
