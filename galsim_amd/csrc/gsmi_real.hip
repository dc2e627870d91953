// Real-space drawing of analytic profiles: drawImage(method='no_pixel' / 'sb'), which the reference
// draws by evaluating each profile at the pixel centres (SBProfile::draw, src/SBProfile.cpp:304-381,
// and the fillXImage loops of src/SBGaussian.cpp, SBExponential.cpp, SBMoffat.cpp, SBSersic.cpp,
// SBBox.cpp).  The reference walks the image row by row, accumulating coordinates, and builds a
// sum of profiles one image at a time; here one thread owns one pixel, evaluates every component of
// the sum there (coordinates from the pixel's indices, components added in the reference's order)
// and writes the pixel once: one pass over the image, 8 bytes written per pixel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/gsmi.h"
#include "device_math.h"
#include "launch.h"

namespace gsmi {
namespace {

struct XComps {
    gsmi_xcomp c[GSMI_XCOMP_MAX];
};

// (1 + r^2)^-beta in the form SBMoffat chose for beta (src/SBMoffat.cpp:210-224, 265-272)
__device__ __forceinline__ double moffat_pow(double x, int kind, double beta)
{
    switch (kind) {
    case 1: return 1. / x;
    case 15: return 1. / (x * sqrt(x));
    case 2: return 1. / (x * x);
    case 25: return 1. / (x * x * sqrt(x));
    case 3: return 1. / (x * x * x);
    case 35: return 1. / (x * x * x * sqrt(x));
    case 4: { const double xsq = x * x; return 1. / (xsq * xsq); }
    default: return exp(-beta * log(x));
    }
}

// The Bessel-function leaves are called out of line: inlined, their register demand would be
// the whole kernel's, for every profile.

// AiryInfoObs / AiryInfoNoObs radial function (src/SBAiry.cpp:88-108, 410-429) at r in units of
// lambda / D; eps = obscuration, thresh = the Taylor threshold on pi r
__device__ __noinline__ double airy_radial(double r, double eps, double thresh)
{
    const double nu = r * M_PI;
    const double xv = (nu < thresh) ? 0.5 * (1. - eps * eps)
                                    : (j1(nu) - eps * j1(eps * nu)) / nu;
    return (xv * xv) * (M_PI / (1. - eps * eps));
}

// SpergelInfo::xValue (src/SBSpergel.cpp:644-651): K_nu(r) r^nu, xnorm0 at r = 0
__device__ __noinline__ double spergel_radial(double r, double nu, double xnorm0)
{
    return (r == 0.) ? xnorm0 : bessel_k(fabs(nu), r) * exp(nu * log(r));
}

// KolmogorovInfo::xValue (src/SBKolmogorov.cpp:114-115): the natural cubic spline of the radial
// table below its last knot, 0 beyond; the knot interval by lower_bound, as Table's lookup of
// unequally spaced arguments
__device__ __noinline__ double kolmogorov_radial(double r, const double* __restrict__ t, int n)
{
    const double* a = t;
    const double* v = t + n;
    const double* y2 = t + 2 * n;
    if (!(r < a[n - 1])) return 0.;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < r) lo = mid + 1; else hi = mid;
    }
    const int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
    const double h = a[i] - a[i - 1];
    const double aa = a[i] - r;
    const double bb = h - aa;
    return (aa * v[i - 1] + bb * v[i] -
            (1. / 6.) * aa * bb * ((aa + h) * y2[i - 1] + (bb + h) * y2[i])) / h;
}

// ceil(v) as an index clamped to [lo, hi] (v may be far outside the image, or NaN: empty range)
__device__ __forceinline__ int clamp_ceil(double v, int lo, int hi)
{
    if (!(v == v)) return hi;
    v = ceil(v);
    return v < double(lo) ? lo : (v > double(hi) ? hi : int(v));
}

// Form-1 Box / TopHat membership of pixel (i, j): 2 strictly inside (|x| < w/2 and |y| < h/2;
// x^2 + y^2 < r^2), 1 on the closed edge, 0 outside.
__device__ __forceinline__ int edge_class(const gsmi_xcomp& c, int i, int j)
{
    const double x = (c.x0 + double(j) * c.dxy) + double(i) * c.dx;
    const double y = (c.y0 + double(j) * c.dy) + double(i) * c.dyx;
    if (c.type == GSMI_LEAF_BOX) {
        if (fabs(x) > c.p[0] || fabs(y) > c.p[1]) return 0;
        return (fabs(x) < c.p[0] && fabs(y) < c.p[1]) ? 2 : 1;
    }
    const double rsq = x * x + y * y;
    if (rsq > c.p[0]) return 0;
    return rsq < c.p[0] ? 2 : 1;
}

// The reference fills a row of a sheared Box or TopHat as zeros while outside, then the norm
// while strictly inside, then zeros (src/SBBox.cpp:136-143, 358-362): if the first pixel of the
// row that is not outside lies exactly on the edge, the whole row stays zero.  Pixel i is
// therefore filled iff it is strictly inside and the pixel before its strictly-inside segment
// (found by bisection: the segment of a convex region along a row is one run) is outside, or the
// segment starts the row.
__device__ bool edge_row_filled(const gsmi_xcomp& c, int i, int j)
{
    if (edge_class(c, i, j) != 2) return false;
    int lo = 0, hi = i;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (edge_class(c, mid, j) == 2) hi = mid; else lo = mid + 1;
    }
    return lo == 0 || edge_class(c, lo - 1, j) == 0;
}

__device__ double xcomp_value(const gsmi_xcomp& c, int i, int j, int m,
                              const double* __restrict__ tables)
{
    double x, y;
    if (c.form == 0) {
        x = c.izero ? double(abs(i - c.izero)) * c.dx : c.x0 + double(i) * c.dx;
        y = c.jzero ? double(abs(j - c.jzero)) * c.dy : c.y0 + double(j) * c.dy;
    } else {
        x = (c.x0 + double(j) * c.dxy) + double(i) * c.dx;
        y = (c.y0 + double(j) * c.dy) + double(i) * c.dyx;
    }
    double v = 0.;
    switch (c.type) {
    case GSMI_LEAF_GAUSSIAN:
        // form 0 is the separable fill: norm * exp(-x^2/2) * exp(-y^2/2)
        v = (c.form == 0) ? (c.norm * exp(-0.5 * x * x)) * exp(-0.5 * y * y)
                          : c.norm * exp(-0.5 * (x * x + y * y));
        break;
    case GSMI_LEAF_EXPONENTIAL:
        v = c.norm * exp(-sqrt(x * x + y * y));
        break;
    case GSMI_LEAF_MOFFAT: {
        const double rsq = x * x + y * y;
        v = (rsq <= c.p[1]) ? c.norm * moffat_pow(1. + rsq, c.kind, c.p[0]) : 0.;
        break;
    }
    case GSMI_LEAF_SERSIC: {
        const double rsq = (i == c.ic && j == c.jc) ? 0. : x * x + y * y;
        v = (c.p[1] > 0. && rsq > c.p[1]) ? 0. : c.norm * exp(-exp(c.p[0] * log(rsq)));
        break;
    }
    case GSMI_LEAF_AIRY:
        v = c.norm * airy_radial(sqrt(x * x + y * y), c.p[0], c.p[1]);
        break;
    case GSMI_LEAF_KOLMOGOROV:
        v = c.norm * kolmogorov_radial(sqrt(x * x + y * y), tables + c.i1, c.i2);
        break;
    case GSMI_LEAF_SPERGEL:
        v = c.norm * spergel_radial(sqrt(x * x + y * y), c.p[0], c.p[1]);
        break;
    case GSMI_LEAF_BOX:
        if (c.form == 0)
            v = (i >= c.i1 && i < c.i2 && j >= c.j1 && j < c.j2) ? c.norm : 0.;
        else
            v = edge_row_filled(c, i, j) ? c.norm : 0.;
        break;
    case GSMI_LEAF_TOPHAT:
        if (c.form == 0) {
            if (j >= c.j1 && j < c.j2) {
                const double yy = c.y0 + double(j) * c.dy;
                const double d = c.p[0] - yy * yy;
                if (d >= 0.) {
                    const double xmax = sqrt(d);
                    const double adx = fabs(c.dx), xr = c.x0 / c.dx;
                    const int a = clamp_ceil(-xmax / adx - xr, 0, m);
                    const int b = clamp_ceil(xmax / adx - xr, 0, m);
                    v = (i >= a && i < b) ? c.norm : 0.;
                }
            }
        } else {
            v = edge_row_filled(c, i, j) ? c.norm : 0.;
        }
        break;
    default:
        break;
    }
    return v * c.amp;
}

// grid: ceil(m n / 256) workgroups (capped), grid-stride over the pixels
__global__ void __launch_bounds__(256)
draw_real_f64(XComps comps, int ncomp, const double* __restrict__ tables, int m, int n,
              int accumulate, double* __restrict__ out)
{
    const size_t total = size_t(m) * size_t(n);
    for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < total;
         p += size_t(gridDim.x) * blockDim.x) {
        const int j = int(p / size_t(m));
        const int i = int(p - size_t(j) * size_t(m));
        double acc = 0.;
        for (int k = 0; k < ncomp; ++k) {
            const double v = xcomp_value(comps.c[k], i, j, m, tables);
            acc = (k == 0) ? v : acc + v;
        }
        out[p] = accumulate ? out[p] + acc : acc;
    }
}

bool xcomp_ok(const gsmi_xcomp& c, const double* tables)
{
    switch (c.type) {
    case GSMI_LEAF_GAUSSIAN: case GSMI_LEAF_EXPONENTIAL: case GSMI_LEAF_MOFFAT:
    case GSMI_LEAF_SERSIC: case GSMI_LEAF_BOX: case GSMI_LEAF_TOPHAT: case GSMI_LEAF_AIRY:
    case GSMI_LEAF_SPERGEL:
        break;
    case GSMI_LEAF_KOLMOGOROV:
        if (!tables || c.i1 < 0 || c.i2 < 2) return false;
        break;
    default:
        return false;
    }
    return c.form == 0 || c.form == 1;
}

}  // namespace
}  // namespace gsmi

extern "C" int gsmi_draw_real_f64(const gsmi_xcomp* comps, int ncomp, const double* tables,
                                  int m, int n, int accumulate, double* out, void* stream)
{
    using namespace gsmi;
    if (!out || m <= 0 || n <= 0 || ncomp < 0 || ncomp > GSMI_XCOMP_MAX ||
        (ncomp > 0 && !comps)) {
        set_error("gsmi_draw_real_f64: invalid argument (need m, n > 0, 0 <= ncomp <= " +
                  std::to_string(GSMI_XCOMP_MAX) + ")");
        return GSMI_E_ARG;
    }
    XComps xc = {};
    for (int k = 0; k < ncomp; ++k) {
        if (!xcomp_ok(comps[k], tables)) {
            set_error("gsmi_draw_real_f64: component " + std::to_string(k) +
                      " has an unsupported type or form (or a Kolmogorov table is missing)");
            return GSMI_E_ARG;
        }
        xc.c[k] = comps[k];
    }
    const size_t total = size_t(m) * size_t(n);
    const size_t want = (total + 255) / 256;
    const int grid = int(want < 8192 ? want : 8192);
    hipLaunchKernelGGL(draw_real_f64, dim3(grid), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), xc, ncomp, tables, m, n,
                       accumulate ? 1 : 0, out);
    return hip_check(hipGetLastError(), "draw_real launch");
}
