/*
 * gsmi -- C ABI of the galsim_amd library (libgsmi.so): GalSim's FFT rendering path,
 * GSObject.drawImage(method='fft'), on AMD Instinct MI355X (gfx950).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device allocations owned by the
 * caller; every launch takes an explicit hipStream_t (passed as void*; 0 = null stream).
 * Every function returns 0 on success or a negative GSMI_E* code; gsmi_last_error() returns a
 * thread-local message for the last failure on the calling thread.  No function needs the
 * Python GIL; the host table caches are guarded by a mutex.
 *
 * Each entry point names the reference interface it replaces (file:line in GalSim 2.8.4).
 */
#ifndef GSMI_H
#define GSMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSMI_OK 0
#define GSMI_E_ARG (-1)       /* invalid argument (shape, size, NaN parameter, ...)      */
#define GSMI_E_RANGE (-2)     /* parameter out of the reference's allowed range           */
#define GSMI_E_SOLVE (-3)     /* a root solve or table build failed                       */
#define GSMI_E_HIP (-4)       /* a HIP runtime call failed                                */
#define GSMI_E_SIZE (-5)      /* FFT size beyond what the kernels support                 */

/* ---------------------------------------------------------------------------------------
 * Lowered stamp description.
 *
 * A profile tree (Transformation / Convolution / Sum over leaves) is lowered on the host into
 * a post-order program evaluated at every k-space sample of the stamp's k-image:
 *   LEAF i    push leaf i evaluated at (kx', ky') = J_i^T (kx, ky)      (SBProfile::drawK)
 *   MUL       pop b, a; push a*b                                        (Convolution)
 *   ADD       pop b, a; push a+b                                        (Sum)
 *   SCALE s   top *= s                                                  (Transformation flux)
 *   PHASE     top *= fs * exp(-i (ux kx + uy ky))                       (ApplyKImagePhases)
 *   INV a b   top = 1 / top where |top| > a, else b (b = 1 / a)         (Deconvolution)
 * A Deconvolution lowers to its original's program, INV, then a KCLIP leaf and MUL: the leaf is
 * 1 where the reference keeps the inverse and 0 where it zeroes it (beyond the original's maxk).
 * ------------------------------------------------------------------------------------- */
enum {
    GSMI_OP_LEAF = 1,
    GSMI_OP_MUL = 2,
    GSMI_OP_ADD = 3,
    GSMI_OP_SCALE = 4,
    GSMI_OP_PHASE = 5,
    GSMI_OP_INV = 6
};

enum {
    GSMI_LEAF_EXPONENTIAL = 1,
    GSMI_LEAF_GAUSSIAN = 2,
    GSMI_LEAF_MOFFAT = 3,
    GSMI_LEAF_KOLMOGOROV = 4,
    GSMI_LEAF_SERSIC = 5,
    GSMI_LEAF_BOX = 6,
    GSMI_LEAF_KCLIP = 7,  /* Deconvolution's maxk cut (galsim/convolve.py:602-612): 0 where
                             ((J00 x + J10 y)^2 + (J01 x + J11 y)^2) dk^2 > maxk^2 at k-image
                             pixel (x, y), else 1; dk in scale, maxk^2 in p[0], in FP64 */
    GSMI_LEAF_SPERGEL = 8, /* flux (1 + (k r0)^2)^(-1-nu), 0 where (k r0)^2 > ksq_max:
                             scale = r0, p[0] = ksq_max, p[1] = -1-nu; cliprows as Exponential */
    GSMI_LEAF_AIRY = 9,   /* knorm A(t), t = (k lam_over_diam / pi)^2, A the overlap area of two
                             unit annuli (inner radius obscuration) at separation sqrt(t)
                             (AiryInfoObs / AiryInfoNoObs::kValue, src/SBAiry.cpp:259-329,
                             409-420): scale = lam_over_diam / pi, flux = knorm =
                             flux / (pi (1 - obs^2)), p[0] = obs, p[1] = obs^2 */
    GSMI_LEAF_INCLINED_EXP = 10, /* kValueHelper of src/SBInclinedExponential.cpp:277-336 at
                             (kx, ky) r0: scale = r0, p[0] = |cos i|,
                             p[1] = pi/2 h0 |sin i| / r0, p[2] = ksq_min, p[3] = ksq_max */
    GSMI_LEAF_INCLINED_VERT = 11, /* InclinedSersic's vertical factor and clip
                             (src/SBInclinedSersic.cpp:321-370): 0 where
                             (kx r0)^2 + (ky r0 cos i)^2 > ksq_max, else x / sinh(x) (Taylor
                             below ksq_min), x = hs ky r0; same fields as INCLINED_EXP */
    GSMI_LEAF_TOPHAT = 12, /* flux 2 J1(k r) / (k r), Taylor below (k r)^2 = 1e-4
                             (SBTopHatImpl::kValue2, src/SBBox.cpp:289-299): scale = radius */
    GSMI_LEAF_DELTA = 13, /* flux everywhere (DeltaFunction, galsim/deltafunction.py) */
    GSMI_LEAF_INTERP_IMAGE = 14 /* SBInterpolatedImage's k value (src/SBInterpolatedImage.cpp:
                             385-448, 767-1009), complex: the k-interpolant's weighted sum of
                             the padded image's stored rfft half-plane at the taps around
                             k N / (2 pi), times the x-interpolant's transform at k / (2 pi).
                             table = offset (doubles) of a GSMI_INTERP_BLOCK-double parameter
                             block (GSMI_INTERP_* indices); ntab = N/2 of the padded image;
                             kind = 1: kValue (no cut but |k| <= maxk1, one tap on an integer
                             k); 0: fillKImage, separable (rows and columns outside the
                             reference's [i1, i2) x [j1, j2) are 0, one tap on an integer k)
                             or, separable = 0, the form with cross terms (0 where |k N /
                             2 pi| > maxk1 N / 2 pi).  Interpolant codes: 0 Nearest,
                             1 Linear, 2 Cubic, 3 Quintic, 4 Lanczos. */
};

typedef struct gsmi_op {
    int32_t code;
    int32_t arg;          /* leaf index for LEAF */
    double a, b, c;       /* SCALE: a; PHASE: ux, uy, fs; INV: threshold, value below it */
} gsmi_op;

/* Parameter block of a GSMI_LEAF_INTERP_IMAGE leaf (doubles in the table buffer). */
#define GSMI_INTERP_N2 0        /* N/2 of the padded image */
#define GSMI_INTERP_KSCALE 1    /* N / (2 pi): k -> k-image index */
#define GSMI_INTERP_MAXK1 2     /* x-interpolant urange * 2 pi */
#define GSMI_INTERP_MAXK1S 3    /* maxk1 * N / (2 pi) */
#define GSMI_INTERP_KXR 4       /* k-interpolant xrange */
#define GSMI_INTERP_I1 5        /* separable fill: first nonzero column (counted from the
                                   k-image's first column) */
#define GSMI_INTERP_I2 6        /* ... one past the last */
#define GSMI_INTERP_J1 7        /* first nonzero row (counted from the k-image's first row) */
#define GSMI_INTERP_J2 8        /* ... one past the last */
#define GSMI_INTERP_XCODE 9     /* x-interpolant code */
#define GSMI_INTERP_XN 10       /* x-interpolant Lanczos order */
#define GSMI_INTERP_XUMAX 11    /* x-interpolant: transform taken as 0 beyond (Lanczos) */
#define GSMI_INTERP_XUTAB 12    /* Lanczos: offset of its transform's spline table (knots,
                                   values, second derivatives), -1 if none */
#define GSMI_INTERP_XUTABN 13   /* number of knots */
#define GSMI_INTERP_KCODE 14    /* k-interpolant code */
#define GSMI_INTERP_KN 15       /* k-interpolant Lanczos order */
#define GSMI_INTERP_KDC 16      /* k-interpolant Lanczos conserve_dc */
#define GSMI_INTERP_KK 17       /* 5 values: its K(1..5) flux-correction coefficients */
#define GSMI_INTERP_KIMAGE 22   /* offset of the k-image: (N rows ky = -N/2..N/2-1) x
                                   (N/2 + 1 columns kx = 0..N/2) complex, interleaved */
#define GSMI_INTERP_XMIN 23     /* the k-image's first column index (kx = index dk) */
#define GSMI_INTERP_YMIN 24     /* ... and first row index */
#define GSMI_INTERP_BLOCK 25
#define GSMI_INTERP_MAX_TAPS 17 /* taps per axis the kernels take (Lanczos order <= 8) */

typedef struct gsmi_leaf {
    int32_t type;         /* GSMI_LEAF_* */
    int32_t separable;    /* 1: diagonal/no Jacobian (the reference's quadrant fill) */
    int32_t kind;         /* Moffat: 15,2,25,3,35,4 or 0; 1 = truncated (kV_trunc table,
                             p[0] = beta, p[1] = 1 / knot spacing in k rD); others 0 */
    int32_t table;        /* Sersic, truncated Moffat: offset of its table in the table buffer
                             (doubles: knots, values, spline second derivatives) */
    int32_t ntab;         /* Sersic, truncated Moffat: number of knots */
    int32_t cliprows;     /* Exponential (FP64): offset of its int2 row ranges, -1 if none;
                             Sersic (FP32): offset (floats, a multiple of 4) in tables32 of its
                             spline as per-interval cubic coefficients (c0, c1, c2, c3) of
                             s = (ln k - knot) / spacing in [0, 1], ntab - 1 quadruples, or -1
                             (then the kernels evaluate the knots' spline form) */
    double jac[4];        /* composite Jacobian (dudx, dudy, dvdx, dvdy) of the leaf       */
    double scale;         /* r0, sigma, rD, 1/k0, r0, or width/2pi                         */
    double scale2;        /* Box: height/2pi                                                */
    double flux;          /* flux (Moffat: knorm)                                            */
    double p[6];          /* type-specific constants (see gsmi_kernels.hip)                  */
} gsmi_leaf;

/* Stack depth of the device program (FP64 / FP32 kernels). */
#define GSMI_STACK_F64 6
#define GSMI_STACK_F32 4

typedef struct gsmi_stamp {
    int32_t N;            /* FFT (wrap) size                                                  */
    int32_t Nk;           /* k-image size before wrapping (Nk >= N)                           */
    int32_t nx, ny;       /* output stamp shape                                               */
    int32_t x0, y0;       /* position of the first output pixel on the N-grid (-N/2 <= x0)    */
    int32_t op_begin;     /* first op of this stamp's program                                 */
    int32_t op_count;     /* ops evaluated per k sample                                       */
    int32_t leaf_begin;   /* leaf indices in LEAF ops are relative to this                    */
    int32_t axis_count;   /* separable factors stored after the program: Box leaves with a
                             diagonal Jacobian, PHASE and SCALE applying to the whole result  */
    int64_t out_offset;   /* element offset of the stamp in the output buffer                 */
    int64_t scratch_offset; /* complex offset in scratch of the stamp's column-transformed
                               values: N/2+1 columns of ny output rows, column-major, columns
                               col_stride apart                                              */
    int64_t axis_offset;  /* complex offset in scratch of its (Nk/2+1) + (Nk+1) axis factors  */
    double dk;            /* k spacing in pixel units: 2 pi / N                               */
    int32_t status;       /* written by the kernels: cleared by each launch's axis prepass,
                             GSMI_STAMP_NONFINITE set by the row transform when a pixel of the
                             stamp is NaN or infinite                                          */
    int32_t col_stride;   /* complex values between the scratch columns: >= ny (0, or less than
                             ny, means ny); a multiple of 16 keeps every column on 128-byte
                             lines for the row transform's reads                               */
} gsmi_stamp;

/* Per-stamp status bits of gsmi_stamp.status. */
#define GSMI_STAMP_NONFINITE 1

/* A work item of the work-list kernels: a group of columns (fold) or rows (row transform) of
 * one stamp.  per_group sizes the workgroup's LDS tile (columns or rows per chunk).  A row item
 * holds count <= per_group rows; a fold item may hold more columns than per_group, which the
 * workgroup takes in chunks of per_group after one set-up of the stamp (program, leaf
 * constants, axis factors and twiddles, staged in LDS when fold_lds has room for them). */
typedef struct gsmi_work {
    int32_t stamp;        /* index into the stamps array                                      */
    int32_t first;        /* first column (c = 0..N/2-1: column 0 also carries kx = N/2, whose
                             transform along ky, like column 0's, is real) or first output row */
    int32_t count;        /* columns or rows in this item                                      */
    int32_t per_group;    /* columns or rows per chunk                                         */
} gsmi_work;

/* A whole batch -- every FFT size -- as two work lists, run by one launch each.  Lists are
 * device arrays; a grid smaller than the list makes the kernel persistent (each workgroup takes
 * items b, b + grid, ...).  lds bytes must cover 2 per_group N (fold) and 2 per_group N/2 (rows)
 * complex values of every item; items that do not fit are skipped.  A fold item stages its
 * stamp's N twiddles and 3 Nk/2 + 2 axis factors in LDS when fold_lds also covers them (after
 * the decoded leaves of a specialised program), and otherwise reads them from memory. */
typedef struct gsmi_schedule {
    const gsmi_work* fold;
    const gsmi_work* rows;
    int32_t nfold, nrows;
    int32_t fold_grid, rows_grid;
    int32_t fold_lds, rows_lds;
    int32_t program;      /* 0, or a program shape compiled into a specialised fold kernel:
                             1 (Exponential + Sersic) x Kolmogorov, 2 Exponential x Moffat,
                             3 Sersic x Kolmogorov, 4 Exponential x Kolmogorov (leaf slots in
                             that order).  Stamps of another shape still draw correctly (the
                             kernel checks every stamp) at the interpreter's speed.  With a
                             nonzero program fold_lds must also hold 4 decoded leaves (80 B
                             each in FP32, 160 B in FP64). */
} gsmi_schedule;

const char* gsmi_last_error(void);
int gsmi_version(void);

/* ---------------------------------------------------------------------------------------
 * Host-side profile scalars and tables (FP64, computed on the CPU; no device needed).
 * gsp: 13 doubles in the order of GSParams(...) (galsim/gsparams.py:126-131), NULL = defaults.
 * ------------------------------------------------------------------------------------- */

/* goodFFTSize (src/Image.cpp:1141-1152; pybind pysrc/Image.cpp:85). */
int gsmi_good_fft_size(int n);

/* SBSersic info (src/SBSersic.cpp:223-537 SersicInfo(n, trunc, gsparams); pysrc/SBSersic.cpp
 * SBSersic ctor): out = {maxK, stepK, HLR, b, ksq_min, ksq_max, highk_a, highk_b, kderiv2,
 * kderiv4, flux fraction} in units of the scale radius; trunc (units of the scale radius) = 0
 * for an untruncated profile; *ntab = number of table knots.  The truncated table's Hankel
 * transforms are integrated here to ~1e-15 of the k = 0 integral on the reference's split
 * points; the reference's adaptive integrator stops at its own tolerance. */
int gsmi_sersic_info(double n, double trunc, const double* gsp, double* out, int* ntab);
/* The Sersic k-space table: knots (log k), values f(k) k^2 and spline second derivatives. */
int gsmi_sersic_table(double n, double trunc, const double* gsp, double* args, double* vals,
                      double* y2, int cap);
/* SersicHLR (src/SBSersic.cpp:665-669; pybind pysrc/SBSersic.cpp). */
int gsmi_sersic_hlr(double n, double flux_fraction, double* out);
/* SersicIntegratedFlux (src/SBSersic.cpp:329-333): fraction of the flux within r (units of the
 * scale radius). */
int gsmi_sersic_integrated_flux(double n, double r, double* out);
/* SersicTruncatedScale (src/SBSersic.cpp:693-775): the scale radius for which hlr encloses half
 * of the flux within trunc. */
int gsmi_sersic_truncated_scale(double n, double hlr, double trunc, double* out);
/* KolmogorovInfo (src/SBKolmogorov.cpp:337-419): out = {maxK, stepK} in units of k0. */
int gsmi_kolmogorov_info(const double* gsp, double* out);
/* SBMoffat untruncated (src/SBMoffat.cpp:163-244, 497-549): out = {maxK, stepK, knorm,
 * knorm2, kind}. */
int gsmi_moffat_info(double beta, double scale_radius, double flux, const double* gsp,
                     double* out);
/* MoffatCalculateScaleRadiusFromHLR, untruncated (src/SBMoffat.cpp:103-126). */
int gsmi_moffat_scale_radius_from_hlr(double hlr, double beta, double* out);
/* MoffatCalculateScaleRadiusFromHLR with a truncation radius (src/SBMoffat.cpp:127-161: Brent
 * from the untruncated value upwards). */
int gsmi_moffat_scale_radius_from_hlr_trunc(double hlr, double trunc, double beta, double* out);
/* SBMoffat truncated (trunc > 0; src/SBMoffat.cpp:163-244, 340-345, 497-626): out = {maxK,
 * stepK, flux factor, half-light radius, number of table knots, knot spacing in k rD}; the
 * table: kV_trunc on args = (k rD)^2 with its natural-spline second derivatives (cap >= knots).
 * The Hankel transforms of the table are integrated here to ~1e-15 of the k = 0 value on the
 * reference's split points; the reference's adaptive integrator stops at its own tolerance. */
int gsmi_moffat_trunc_info(double beta, double scale_radius, double trunc, const double* gsp,
                           double* out);
int gsmi_moffat_trunc_table(double beta, double scale_radius, double trunc, const double* gsp,
                            double* args, double* vals, double* y2, int cap);
/* Exponential k-range clipping of the drawFFT k-image (bounds 0..Nk/2, -Nk/2..Nk/2, spacing dk)
 * for a leaf of scale radius r0 drawn with Jacobian jac (NULL = identity): out[2j], out[2j+1] =
 * [i1, i2) of row j (ky index j - Nk/2), as SBExponentialImpl::fillKImage computes them through
 * GetKValueRange1d/2d (src/SBProfile.cpp:508-588, src/SBExponential.cpp:442-510). */
int gsmi_exponential_cliprows(double r0, const double* gsp, const double* jac, int Nk, double dk,
                              int32_t* out);

/* The same row ranges for any radial leaf clipped at (k r0)^2 > ksq_max (Spergel: ksq_max =
 * kvalue_accuracy^(-1/(nu+1)) - 1, SBSpergelImpl::fillKImage src/SBSpergel.cpp:456-527). */
int gsmi_kclip_rows(double ksq_max, double r0, const double* jac, int Nk, double dk,
                    int32_t* out);

/* Spergel profile scalars (SpergelInfo, src/SBSpergel.cpp:529-653; pybind SBSpergel and
 * SpergelCalculateHLR, pysrc/SBSpergel.cpp): out = {maxk, stepk, hlr, ksq_max} in units of the
 * scale radius.  -0.85 <= nu <= 4 (GSMI_E_RANGE otherwise). */
int gsmi_spergel_info(double nu, const double* gsp, double* out);
/* InclinedExponential scalars (SBInclinedExponentialImpl ctor,
 * src/SBInclinedExponential.cpp:64-160; pybind pysrc/SBInclinedExponential.cpp) for
 * cosi = |cos i| and hs = pi/2 h0 |sin i| / r0: out = {maxk, stepk, ksq_min, ksq_max}, sizes in
 * units of the scale radius. */
int gsmi_inclined_exp_info(double cosi, double hs, const double* gsp, double* out);
/* kValueHelper (src/SBInclinedExponential.cpp:277-336) at (kx, ky) in units of 1/r0, on the
 * host: out[i] = value / flux. */
int gsmi_inclined_exp_kvalue(double cosi, double hs, double ksq_min, double ksq_max,
                             const double* kx, const double* ky, double* out, int count);

/* InclinedSersic scalars (SBInclinedSersicImpl ctor, src/SBInclinedSersic.cpp:86-209) for
 * Sersic index n, truncation radius trunc (units of r0; 0 = none), cosi = |cos i| and
 * hs = pi/2 h0 |sin i| / r0: out = {maxk (1/r0), ksq_min, ksq_max}.  The profile is lowered as a
 * Sersic leaf whose Jacobian scales ky' by cos i times a GSMI_LEAF_INCLINED_VERT leaf. */
int gsmi_inclined_sersic_info(double n, double trunc, double cosi, double hs, const double* gsp,
                              double* out);

/* SBSpergel::calculateFluxRadius / calculateIntegratedFlux in units of the scale radius. */
int gsmi_spergel_flux_radius(double nu, double f, double* out);
int gsmi_spergel_integrated_flux(double nu, double u, double* out);
/* Host evaluation of a Sersic leaf k-value (SersicInfo::kValue, src/SBSersic.cpp:363-376). */
int gsmi_sersic_kvalue(double n, double trunc, const double* gsp, const double* ksq, double* out,
                       int count);

/* Sine integral Si(x) of n values (math::Si, src/math/Sinc.cpp:101-170; here its power series
 * below |x| = 4 and E1's continued fraction above). */
int gsmi_sine_integral(const double* x, double* out, int n);
/* Lanczos(n, conserve_dc) interpolant (Lanczos::Lanczos, src/Interpolant.cpp:787-895):
 * out = {K(1), .., K(5), umax}, *ntab = knots of its transform's spline table; the table itself
 * (knots u = 0, du, ..., values uCalc(u), spline second derivatives) by gsmi_lanczos_utable into
 * buffers of cap values. */
int gsmi_lanczos_info(int n, int conserve_dc, const double* gsp, double* out, int* ntab);
int gsmi_lanczos_utable(int n, int conserve_dc, const double* gsp, double* args, double* vals,
                        double* y2, int cap);

/* ---------------------------------------------------------------------------------------
 * Device rendering.  All buffers are device pointers; stamps/ops/leaves/tables are device
 * copies of the lowered description; scratch holds sum(col_stride * (N/2+1)) complex values.
 * ------------------------------------------------------------------------------------- */

/* drawImage(method='fft') in FP64 for a batch of stamps that share one FFT size N (max_ny =
 * tallest output stamp; scratch holds each stamp's col_stride * (N/2+1) values (column-major) and its
 * (Nk/2+1) + (Nk+1) axis factors at the offsets in its record, complex128) (the
 * reference: gsobject.py:1927-2045 drawFFT_makeKImage + _drawKImage + drawFFT_finish, with
 * SBProfile::drawK src/SBProfile.cpp:383-425, ApplyKImagePhases src/SBTransform.cpp:763-883,
 * wrapImage src/Image.cpp:577-726 and irfft src/Image.cpp:814-911).  out is float64. */
int gsmi_draw_fft_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny,
                      const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                      const int32_t* cliprows, void* scratch, double* out, void* stream);

/* The same in FP32 arithmetic (FP64 kept for the host tables, the Moffat Bessel K and the flux
 * sums); tables32 is the Sersic table rounded to float; out is float32.  Batched: thousands of
 * stamps of one FFT size per call.  max_ny = tallest output stamp. */
int gsmi_draw_fft_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny,
                      const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                      const float* tables32, void* scratch, float* out, void* stream);

/* The two kernels of gsmi_draw_fft_* as separate launches (same arguments): the fused k-space
 * evaluation + product/sinc/phase + wrap + column inverse DFT into scratch, and the row real
 * inverse DFT + crop + 1/N^2 into out.  Calling both in order equals gsmi_draw_fft_*. */
int gsmi_fold_colfft_f64(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
                         const gsmi_leaf* leaves, const double* tables, const int32_t* cliprows,
                         void* scratch, void* stream);
int gsmi_fold_colfft_f32(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
                         const gsmi_leaf* leaves, const double* tables, const float* tables32,
                         void* scratch, void* stream);
int gsmi_row_c2r_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny,
                     const void* scratch, double* out, void* stream);
int gsmi_row_c2r_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny,
                     const void* scratch, float* out, void* stream);

/* The same drawing as gsmi_draw_fft_* for a batch of any mix of FFT sizes, scheduled by work
 * lists (see gsmi_schedule): the axis-factor prepass over all nstamps, one fold launch and one
 * row launch.  Scratch and output offsets are those of the stamp records.  gsmi_fold_list_* /
 * gsmi_rows_list_* are the two halves (the prepass runs with the fold). */
int gsmi_draw_list_f64(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const int32_t* cliprows, void* scratch, double* out, void* stream);
int gsmi_draw_list_f32(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const float* tables32, void* scratch, float* out, void* stream);
int gsmi_fold_list_f64(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const int32_t* cliprows, void* scratch, void* stream);
int gsmi_fold_list_f32(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const float* tables32, void* scratch, void* stream);
int gsmi_rows_list_f64(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const void* scratch, double* out, void* stream);
int gsmi_rows_list_f32(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const void* scratch, float* out, void* stream);

/* Evaluate a stamp program on a k-image with bounds (xmin..xmin+nx-1, ymin..ymin+ny-1) and
 * spacing dk, writing complex128 values row-major (SBProfile::drawK src/SBProfile.cpp:383-425
 * through the Python _drawKImage tree, gsobject.py:2464-2583 drawKImage). */
int gsmi_draw_kimage_f64(const gsmi_op* ops, int nops, const gsmi_leaf* leaves,
                         const double* tables, const int32_t* cliprows, int xmin, int ymin,
                         int nx, int ny, int Nk, double dk, double* out, void* stream);

/* irfft(in, out, shift_in=true, shift_out=true) (src/Image.cpp:814-911; pybind
 * pysrc/Image.cpp:76): in is the wrapped (N/2+1) x N complex128 k-image, rows ky=-N/2..N/2-1
 * (row-major, x fastest); out is the N x N float64 real image on (-N/2..N/2-1)^2; scratch holds
 * N * (N/2+1) complex128. */
int gsmi_irfft_f64(const double* kin, int N, double* out, void* scratch, void* stream);

/* gsmi_irfft_f64 without allocating: the same result, with its N (N/2+1) complex128 scratch and
 * twiddle tables in a static device array of the library (about 33.7 MB per device, held until
 * the process ends) instead of a caller buffer.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture after a first call on the device,
 * tight serving loops, harnesses that forbid hipMalloc).  N <= GSMI_NOALLOC_MAX_N, else
 * GSMI_E_SIZE.  The first call on each device uploads the twiddle tables with synchronous
 * copies (make it outside any capture).  Calls that overlap on one device race on the shared
 * array. */
#define GSMI_NOALLOC_MAX_N 2048
int gsmi_irfft_f64_noalloc(const double* kin, int N, double* out, void* stream);

/* ApplyKImagePhases (src/SBTransform.cpp:763-883; pybind pysrc/SBTransform.cpp:34-46) on a
 * device k-image: nx x ny complex128, row-major, pixel (xmin + i, ymin + j) at k = (x, y) dk,
 * multiplied by flux_scaling exp(-i (kx' cenx + ky' ceny)) with (kx', ky') the k mapped by the
 * host 2x2 Jacobian jac (dudx, dudy, dvdx, dvdy; null for none) as in SBProfile::drawK.  The
 * phase is evaluated per pixel (the reference's renormalised recurrence drifts by ~N ulp). */
int gsmi_apply_kimage_phases_f64(double* kimage, int xmin, int ymin, int nx, int ny, double dk,
                                 const double* jac, double cenx, double ceny, double flux_scaling,
                                 void* stream);

/* wrapImage(hermx=true, hermy=false) (src/Image.cpp:577-726; pybind pysrc/Image.cpp:80):
 * folds the (Nk/2+1) x (Nk+1) k-image (rows ky=-Nk/2..Nk/2) onto (N/2+1) x N (rows
 * ky=-N/2..N/2-1).  Both complex128 row-major. */
int gsmi_wrap_hermx_f64(const double* kin, int Nk, int N, double* kout, void* stream);

/* ---------------------------------------------------------------------------------------
 * Two-dimensional DFTs of M x N images, M and N even (galsim.fft fft2 / ifft2 / rfft2 / irfft2,
 * galsim/fft.py:29-300, and Image.calculate_fft / calculate_inverse_fft, galsim/image.py:915-1016).
 * All buffers are device memory, row-major (x fastest), complex values as interleaved
 * (re, im) float64.  Results equal numpy.fft's: with shift_in the input is fftshift-ed first,
 * with shift_out the output is fftshift-ed after (only along y for the half-plane output of
 * rfft2).  Line lengths with prime factors 2, 3, 5 only are transformed directly; other lengths
 * L by Bluestein's algorithm with a 5-smooth length >= 2L - 1, which must fit the workgroup's
 * LDS; a column length M beyond that takes a four-step transform (M = M1 M2, both even and
 * transformable in LDS); GSMI_E_SIZE otherwise.  Work is queued on `stream`; no call allocates device memory
 * except the per-length twiddle / chirp tables, made once per device.
 * --------------------------------------------------------------------------------------- */

/* cfft(in, out, inverse, shift_in, shift_out) (src/Image.cpp:913-990; pybind
 * pysrc/Image.cpp:59-63): in and out are M x N complex; the inverse carries 1/(M N).  scratch
 * holds M x N complex.  out may equal in. */
int gsmi_cfft2_f64(const double* in, int M, int N, int inverse, int shift_in, int shift_out,
                   double* out, void* scratch, void* stream);

/* rfft(in, out, shift_in, shift_out) (src/Image.cpp:729-812; pybind pysrc/Image.cpp:56,61): in is
 * M x N real, out M x (N/2+1) complex (kx = 0..N/2); scratch holds M x (N/2+1) complex. */
int gsmi_rfft2_f64(const double* in, int M, int N, int shift_in, int shift_out, double* out,
                   void* scratch, void* stream);

/* irfft(in, out, shift_in, shift_out) (src/Image.cpp:814-911; pybind pysrc/Image.cpp:58,62) for
 * any even M x N: in is M x (N/2+1) complex, out M x N real, with 1/(M N); the imaginary parts of
 * the kx = 0 and kx = N/2 columns after the y transform are ignored.  scratch holds
 * M x (N/2+1) complex, twice that when M takes the four-step column transform
 * (gsmi_fft_column_passes(M) == 2). */
int gsmi_irfft2_f64(const double* in, int M, int N, int shift_in, int shift_out, double* out,
                    void* scratch, void* stream);

/* 1 if columns of length M transform in one LDS pass on the current device, 2 if they take the
 * four-step transform, or a negative GSMI_E_* code if M is not supported.  No counterpart in the
 * reference (FFTW plans any length). */
int gsmi_fft_column_passes(int M);

/* The M x N transform of a real image from its M x (N/2+1) rfft2 half-plane (with the same
 * shift_out), by Hermitian symmetry: the fill in galsim/fft.py:88-103. */
int gsmi_hermitian_fill_f64(const double* half, int M, int N, int shift_out, double* out,
                            void* stream);

/* ---------------------------------------------------------------------------------------
 * Adjoint of gsmi_draw_fft_* (no counterpart in the reference, which has no derivatives: it is
 * the exact reverse-mode derivative of drawFFT, galsim/gsobject.py:2014-2045, as restated by
 * the two kernels above).
 *
 * Given resid = dL/dI for every output pixel (same layout and precision as out of
 * gsmi_draw_fft_*), writes grad[s * GSMI_GRAD_FIELDS + f] = dL/dfield_f of stamp s (stamps in
 * the order of the stamps array), FP64:
 *   f = 7 l + {0..6}: leaf slot l (the relative leaf index of a LEAF op of the program, l <
 *                     GSMI_GRAD_MAXLEAF): flux, scale, jac[0], jac[1], jac[2], jac[3], and the
 *                     shape p[0] -- for a Moffat leaf its beta, the derivative of the Moffat
 *                     k-profile in beta at fixed knorm and scale (GSMI_ADJ_SHAPE only);
 *   f = 28, 29:       ux, uy of a PHASE applying to the whole result (hoisted axis factor);
 *   f = 30:           log fs of a SCALE / PHASE amplitude applying to the whole result.
 * Only radial leaves (Exponential, Gaussian, Moffat, Kolmogorov, Sersic) have field gradients;
 * a Box slot reads zero, and so does the shape field of every leaf but Moffat.  Exponential clip
 * rows (FP64) are held fixed: clipped samples have zero derivative, as they have zero value.
 * max_leaves bounds the number of leaf slots the programs use (1..GSMI_GRAD_MAXLEAF).  flags:
 * GSMI_ADJ_COMPLEX if any program holds a PHASE op (a shift inside the tree rather than on the
 * whole result), GSMI_ADJ_SHAPE to compute the Moffat beta derivatives (three FP64 Bessel
 * evaluations per Moffat sample), and in bits 8-15 a program shape of gsmi_schedule.program
 * that every stamp has (a specialised kernel; 0 = the interpreter).  A stamp whose program
 * breaks a promise, or has more than 12 ops, gets NaN gradients.  scratch is the forward's scratch layout (it is overwritten);
 * partials holds gsmi_adjoint_partials() doubles.
 * ------------------------------------------------------------------------------------- */
#define GSMI_GRAD_MAXLEAF 4
#define GSMI_GRAD_PER_LEAF 7
#define GSMI_GRAD_FIELDS 31
#define GSMI_ADJ_COMPLEX 1
#define GSMI_ADJ_SHAPE 2

int gsmi_adjoint_partials(int N, int nstamps, int64_t* ndoubles);
int gsmi_adjoint_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, int max_leaves,
                     int flags, const gsmi_op* ops, const gsmi_leaf* leaves,
                     const double* tables, const int32_t* cliprows, const double* resid,
                     void* scratch, double* partials, double* grad, void* stream);
int gsmi_adjoint_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, int max_leaves,
                     int flags, const gsmi_op* ops, const gsmi_leaf* leaves,
                     const double* tables, const float* tables32, const float* resid,
                     void* scratch, double* partials, double* grad, void* stream);

/* -------------------------------------------------------------------------------------------
 * Parameter map of the differentiable Exponential (x) Moffat model (no counterpart in the
 * reference: it restates on the device what the host lowering does with galsim/exponential.py,
 * galsim/shear.py:365-386 and galsim/moffat.py, so a fit step needs no host work).  theta rows
 * (n x 7, or n x 8 with fit_beta): flux, hlr, g1, g2, dx, dy, psf_fwhm [, beta], FP64 device
 * arrays.
 *   gsmi_exp_moffat_fields: the GSMI_GRAD_FIELDS fields of each stamp (base[i] where no
 *     parameter reaches; cxcy[i] the centring shift added to (dx, dy); fwhm_to_rd[i] the fixed
 *     rD / fwhm without fit_beta) written into the device records at the double indices
 *     dst[i * GSMI_GRAD_FIELDS + f] (< 0: not written), and with fit_beta the 5 Moffat values
 *     that follow beta (the k = 0 value and Temme's constants) at aux_dst[i * 5 + k].
 *   gsmi_exp_moffat_grad: dL/dtheta (n x 7 or 8) from dL/dfields (gsmi_adjoint_* output, row
 *     inv[i] for stamp i).
 * ------------------------------------------------------------------------------------- */
int gsmi_exp_moffat_fields(const double* theta, int n, int fit_beta, const double* base,
                           const double* cxcy, double inv_scale, const double* fwhm_to_rd,
                           const int64_t* dst, const int64_t* aux_dst, double* records,
                           void* stream);
int gsmi_exp_moffat_grad(const double* theta, int n, int fit_beta, const double* fwhm_to_rd,
                         const double* gfields, const int64_t* inv, double inv_scale,
                         double* dtheta, void* stream);

/* -------------------------------------------------------------------------------------------
 * Real-space drawing: drawImage(method='no_pixel' / 'sb') of analytic profiles, which the
 * reference draws by direct evaluation at pixel centres (GSObject.drawReal / _drawReal,
 * galsim/gsobject.py:1846-1898, galsim/transform.py:571-583, galsim/sum.py:304-310, into
 * SBProfile::draw, src/SBProfile.cpp:304-381, and each profile's fillXImage).  One component per
 * transformed leaf; the host has applied draw()'s offset / Jacobian set-up and the leaf's own
 * radius scaling, so pixel (i, j) of an m x n image sits at
 *   form 0: x = x0 + i dx, y = y0 + j dy, or |i - izero| dx, |j - jzero| dy when izero / jzero
 *           are nonzero (the quadrant fill, src/SBProfile.cpp:447-488, for the radial leaves);
 *   form 1: x = x0 + i dx + j dxy, y = y0 + i dyx + j dy.
 * The pixel value is sum_c amp_c * (norm_c f_c(x, y)), components added in order, with f_c:
 *   GAUSSIAN   exp(-(x^2 + y^2) / 2) (form 0: exp(-x^2/2) exp(-y^2/2), src/SBGaussian.cpp:120-190)
 *   EXPONENTIAL exp(-sqrt(x^2 + y^2))                           (src/SBExponential.cpp:381-440)
 *   MOFFAT     (1 + r^2)^-beta for r^2 <= p[1], else 0; kind selects 1/x, 1/x^1.5, ... 1/x^4
 *              (1, 15, 2, 25, 3, 35, 4) or exp(-beta log x) (0), p[0] = beta
 *                                                         (src/SBMoffat.cpp:210-272, 340-415)
 *   SERSIC     exp(-exp(p[0] log r^2)) (p[0] = 1 / 2n), 0 where p[1] > 0 and r^2 > p[1]; pixel
 *              (ic, jc) >= 0 takes r = 0                  (src/SBSersic.cpp:104-222, 354-360)
 *   AIRY       nu = pi r: (pi / (1 - e^2)) ((J1(nu) - e J1(e nu)) / nu)^2, or with
 *              (1 - e^2) / 2 inside the bracket for nu < p[1]; e = p[0]   (src/SBAiry.cpp:88-190)
 *   KOLMOGOROV the spline of gsmi_kolmogorov_xtable at r = |(x, y)| below its last knot, else 0
 *                                                  (src/SBKolmogorov.cpp:104-115, 127-185)
 *   SPERGEL    K_nu(r) r^nu (nu = p[0]), p[1] at r = 0 (infinite for nu <= 0)
 *                                                         (src/SBSpergel.cpp:80-108, 394-452, 644-651)
 *   BOX        1 inside; form 0: i1 <= i < i2 and j1 <= j < j2; form 1: |x| < p[0], |y| < p[1]
 *                                                                (src/SBBox.cpp:80-148)
 *   TOPHAT     1 inside; form 0: rows j1 <= j < j2, columns from xmax = sqrt(p[0] - y^2);
 *              form 1: x^2 + y^2 < p[0]                          (src/SBBox.cpp:301-370)
 * ------------------------------------------------------------------------------------- */
#define GSMI_XCOMP_MAX 16

typedef struct gsmi_xcomp {
    int32_t type;         /* GSMI_LEAF_GAUSSIAN, _EXPONENTIAL, _MOFFAT, _KOLMOGOROV, _SERSIC,
                             _AIRY, _SPERGEL, _BOX, _TOPHAT */
    int32_t form;         /* 0 or 1, as above */
    int32_t kind;         /* Moffat power form */
    int32_t izero, jzero; /* form 0 quadrant origin */
    int32_t ic, jc;       /* Sersic centre pixel (-1: none) */
    int32_t i1, i2, j1, j2; /* Box (form 0) index ranges; TopHat rows j1..j2; Kolmogorov: table
                               offset i1 and knots i2 */
    int32_t pad;
    double x0, dx, dxy, y0, dy, dyx;
    double norm;          /* the leaf's amplitude */
    double amp;           /* draw()'s flux ratio, applied after the fill */
    double p[2];
} gsmi_xcomp;

/* Fill the m x n FP64 image `out` (device, row-major, row j = y) with the components comps[0..
 * ncomp) (a host array, ncomp <= GSMI_XCOMP_MAX), adding to its contents when accumulate is 1.
 * `tables` (device, may be NULL without Kolmogorov components) holds each Kolmogorov component's
 * radial table at tables + i1: i2 knots, then i2 values, then i2 second derivatives.
 * Replaces the fillXImage loops of SBProfile::draw (src/SBProfile.cpp:304-381). */
int gsmi_draw_real_f64(const gsmi_xcomp* comps, int ncomp, const double* tables, int m, int n,
                       int accumulate, double* out, void* stream);

/* KolmogorovInfo's x-space radial table (src/SBKolmogorov.cpp:337-384: f(0) then the Hankel
 * transform of exp(-k^(5/3)) / 2 pi on log r from -3, natural cubic spline): returns the number
 * of knots, and fills args / vals / y2 (second derivatives) when n is at least that. */
int gsmi_kolmogorov_xtable(const double* gsp, double* args, double* vals, double* y2, int n);

#ifdef __cplusplus
}
#endif
#endif
