// Adjoint of the FFT drawing path: the gradient of a pixel loss with respect to the fields of the
// lowered stamp description, given the loss gradient dL/dI at every output pixel.
//
// The forward path computes, per stamp (p = pixel position, k-image samples s = (i, j) with
// kx = i dk, ky = j dk, wrapped cells (c, r) with c = 0..N/2):
//   I(p) = 1/N^2 sum_{c,r} w_c Re[ W(c, r) e^{+2 pi i (c px + r py) / N} ],  w_c = 1 for c = 0
//          or N/2 and 2 otherwise,
//   W(c, r) = sum_{s = (c, r) mod N} F(s) + sum_{s = -(c, r) mod N, i > 0} conj F(s),
//   F(s) = P(s) X(i) Y(j)  (program value times the hoisted axis factors).
// So for any field theta,
//   dL/dtheta = sum_{c,r} Re[ dW(c, r)/dtheta A(c, r) ],  A = w_c Rhat(c, r) / N^2,
//   Rhat(c, r) = sum_p R(p) e^{+2 pi i (c px + r py) / N},  R = dL/dI,
// and each aliased sample contributes Re[dF(s)/dtheta A] (or Re[dF(s)/dtheta conj A] where the
// forward added conj F).  The kernels mirror the forward ones in reverse order:
//   1. row_r2c: per output row, the real DFT (sign +) of the residual row, placed on the N-grid,
//      through a half-length complex DFT in LDS; writes the (N/2+1) coefficients to scratch (the
//      same rows the forward's fold kernel writes).
//   2. grad_colfft: per group of columns c, the column DFT of those rows in LDS gives Rhat(c, .);
//      each thread then walks the same aliased samples as the forward, re-evaluates the leaves
//      with their q-derivatives (no complex grid is stored between forward and backward), gets
//      dP/dv_l of the post-order program from P(v_l = 1) - P(v_l = 0) (P is affine in every
//      leaf value), and accumulates the field gradients; a block reduction writes one partial
//      vector per workgroup.
//   3. grad_reduce: per stamp, sums the workgroup partials in a fixed order (deterministic).
//
// Fields, per stamp (GSMI_GRAD_FIELDS doubles): for each leaf slot l < GSMI_GRAD_MAXLEAF
// (relative leaf index of the program's LEAF ops) d/d(flux, scale, jac[0..3], shape) at
// 7 l .. 7 l + 6, then d/dux, d/duy of a phase applying to the whole result and d/dlog(fs) of a
// factor applying to the whole result.  The shape field is the Moffat beta.  Radial leaves only (Exponential, Gaussian, Moffat, Kolmogorov, Sersic):
// a Box leaf in the program contributes its value, and its slot reads zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/gsmi.h"
#include "device_program.h"
#include "launch.h"

namespace gsmi {

namespace {

// Unit-flux leaf value phi and dphi/dq, q = scale^2 |J^T k|^2 (radial leaves), at sample (i, j).
template <typename T>
struct LeafD {
    T phi;
    T dphi;
};

// Moffat Bessel kinds in FP64: g(k) and g'(k) / (2 k) for k > 0 (d/dk [k^nu K_nu(k)] =
// -k^nu K_{nu-1}(k)).
__device__ __noinline__ LeafD<double> moffat_bessel_d(int kind, double nu, double k, double ksq,
                                                      double at_zero)
{
    if (ksq == 0.)
        return {(kind == 2) ? 1. : (kind == 3) ? 2. : (kind == 4) ? 8. : at_zero, 0.};
    double km1;
    switch (kind) {
      case 2: {
        double k1 = bessel_k(1., k, &km1);
        return {k1 * k, -0.5 * km1};
      }
      case 3: {
        double k2 = bessel_k(2., k, &km1);
        return {k2 * ksq, -0.5 * k * km1};
      }
      case 4: {
        double k3 = bessel_k(3., k, &km1);
        return {k3 * k * ksq, -0.5 * ksq * km1};
      }
      default: {
        double kn = bessel_k(nu, k, &km1);
        double pk = pow(k, nu - 1.);
        return {kn * pk * k, -0.5 * pk * km1};
      }
    }
}

// The forward's eval_leafc at unit flux, with dphi/dq.  The branches and clipping follow
// eval_leafc exactly, so phi * flux is the forward value.
// d/dnu [k^nu K_nu(k)] (nu = beta - 1 >= 0.1) in FP64: k^nu (ln k K_nu + dK_nu/dnu), the order
// derivative by a central difference with step 1e-4 (truncation ~1e-9 relative, rounding
// ~1e-12); at k = 0 the limit 2^(nu-1) Gamma(nu) (ln 2 + psi(nu)), psi by the difference of
// lgamma.
__device__ __noinline__ double moffat_dg_dnu(double nu, double k)
{
    const double h = 1e-4;
    if (k == 0.) {
        const double g0 = exp((nu - 1.) * M_LN2 + lgamma(nu));
        const double psi = (lgamma(nu + h) - lgamma(nu - h)) / (2. * h);
        return g0 * (M_LN2 + psi);
    }
    const double lk = log(k);
    const double kp = bessel_k(nu + h, k);
    const double km = bessel_k(nu - h, k);
    const double kn = bessel_k(nu, k);
    return exp(nu * lk) * (lk * kn + (kp - km) / (2. * h));
}

// The same derivative in FP32 from the integral K_nu(k) = int_0^inf exp(-k cosh t) cosh(nu t) dt
// (Abramowitz & Stegun 9.6.24), whose nu-derivative is int_0^inf t sinh(nu t) exp(-k cosh t) dt.
// With L = ln k both become sums over nodes t_m of e+ = exp(nu (t_m + L) - k cosh t_m) and
// e- = exp(nu (L - t_m) - k cosh t_m):  k^nu K_nu = sum w_m (e+ + e-) / 2 and
// d/dnu [k^nu K_nu] = sum w_m ((t_m + L) e+ + (L - t_m) e-) / 2, the trapezoid rule with step
// h = 11/32.  The integrands are analytic in |Im t| < pi/2 and decay double-exponentially, so
// the rule converges like exp(-pi^2 / h); measured against FP64 Bessel functions its error is
// below 1.2e-6 of the k = 0 value for nu <= 5.5 (3e-6 at nu = 8), over k in [1e-4, 60].  The sum
// stops once e+ is past its peak and below 1e-9 of the running sum (e- <= e+ throughout); at
// most 48 nodes (t <= 16.2, enough for k >= 1e-5).  L is split as e ln2_hi + (ln m + e ln2_lo)
// (k = m 2^e, m in [1/sqrt2, sqrt2)), so t_m + e ln2_hi is exact and L carries no rounding of
// its large part into the exponents.  k = 0 takes the FP64 limit.
__device__ inline float moffat_dg_dnu_f32(float nu, float k)
{
    if (k == 0.f) return float(moffat_dg_dnu(double(nu), 0.));
    constexpr float kCosh[48] = {
        1.000000000e+00f, 1.059666109e+00f, 1.245784524e+00f, 1.580565168e+00f, 2.103958159e+00f,
        2.878421143e+00f, 3.996372503e+00f, 5.591219856e+00f, 7.853279873e+00f, 1.105248919e+01f,
        1.557061655e+01f, 2.194682011e+01f, 3.094198637e+01f, 4.362952848e+01f, 6.152347897e+01f,
        8.675916262e+01f, 1.223480095e+02f, 1.725369157e+02f, 2.433150346e+02f, 3.431284761e+02f,
        4.838881994e+02f, 6.823913746e+02f, 9.623258256e+02f, 1.357096751e+03f, 1.913813041e+03f,
        2.698908885e+03f, 3.806071511e+03f, 5.367421090e+03f, 7.569276928e+03f, 1.067439137e+04f,
        1.505330460e+04f, 2.122856204e+04f, 2.993707087e+04f, 4.221803674e+04f, 5.953697455e+04f,
        8.396059155e+04f, 1.184034121e+05f, 1.669755744e+05f, 2.354733022e+05f, 3.320705813e+05f,
        4.682945791e+05f, 6.604012075e+05f, 9.313149763e+05f, 1.313364626e+06f, 1.852140989e+06f,
        2.611937444e+06f, 3.683422184e+06f, 5.194457862e+06f,
    };
    constexpr float h = 0.34375f;
    constexpr float kLog2e = 1.44269504088896341f;
    int ex;
    float mant = frexpf(k, &ex);
    if (mant < 0.70710678f) {
        mant *= 2.f;
        --ex;
    }
    const float eh = float(ex) * 0.693145751953125f;       // 12-bit constant: exact
    const float el = __logf(mant) + float(ex) * 1.428606765330187e-6f;
    const float nu2 = nu * kLog2e;
    const float k2 = k * kLog2e;
    float S = 0.f, D = 0.f, eprev = 0.f;
#pragma unroll
    for (int m = 0; m < 48; ++m) {
        const float t = float(m) * h;
        const float w = (m == 0) ? 0.25f * h : 0.5f * h;
        const float up = (t + eh) + el;
        const float um = (eh - t) + el;
        const float ep = __builtin_amdgcn_exp2f(nu2 * up - k2 * kCosh[m]);
        const float em = __builtin_amdgcn_exp2f(nu2 * um - k2 * kCosh[m]);
        S += w * (ep + em);
        D += w * (up * ep + um * em);
        if (ep < eprev && ep < 1e-9f * S) break;
        eprev = ep;
    }
    return D;
}

template <typename T>
__device__ inline LeafD<T> eval_leafc_d(const LeafC<T>& C, const LeafI& I, const gsmi_leaf* rec,
                                        const Tables<T>& tb, T kx, T ky, int i, int j, int Nk)
{
    T kxp, kyp;
    leaf_kprime(C, I, kx, ky, &kxp, &kyp);
    const T X = kxp * C.s;
    const T Y = kyp * C.s;
    const T ksq = X * X + Y * Y;
    switch (I.type) {
      case GSMI_LEAF_EXPONENTIAL: {
        if (exp_clip<T>(C, I, tb, i, j, Nk, ksq)) return {T(0), T(0)};
        T q = T(1) + ksq;
        if constexpr (sizeof(T) == 4) {
            const float r = __builtin_amdgcn_rsqf(q);         // q^(-1/2)
            const float phi = r * r * r;
            return {phi, -1.5f * phi * (r * r)};
        }
        T phi = ddiv(T(1), q * dsqrt(q));
        return {phi, ddiv(T(-1.5) * phi, q)};
      }
      case GSMI_LEAF_GAUSSIAN: {
        if (I.sep) {
            T phi = dexp(T(-0.5) * X * X) * dexp(T(-0.5) * Y * Y);
            return {phi, T(-0.5) * phi};
        }
        if (ksq > C.p[0]) return {T(0), T(0)};
        if (ksq < C.p[1])
            return {T(1) - T(0.5) * ksq * (T(1) - T(0.25) * ksq), T(-0.5) + T(0.25) * ksq};
        T phi = dexp(T(-0.5) * ksq);
        return {phi, T(-0.5) * phi};
      }
      case GSMI_LEAF_MOFFAT: {
        T k;
        if constexpr (sizeof(T) == 4) k = __builtin_amdgcn_sqrtf(ksq);
        else k = dsqrt(ksq);
        T e = dexp(-k);
        switch (I.kind) {
          case 1: {
            T dS;
            const T S = moffat_trunc_spline<T>(sersic_table<T>(tb, I.aux), I.ntab, ksq, k,
                                               C.p[1], &dS);
            return {S, dS};
          }
          case 15:
            return {e, ksq == T(0) ? T(0) : ddiv(T(-0.5) * e, k)};
          case 25:
            return {e * (T(1) + k), T(-0.5) * e};
          case 35:
            return {e * (T(3) + (T(3) + k) * k), T(-0.5) * e * (T(1) + k)};
          default: {
            if constexpr (sizeof(T) == 4) {
                if (I.kind == 2 || I.kind == 3 || I.kind == 4) {
                    float g, dg;
                    moffat_int_kind_f32(I.kind, k, &g, &dg);
                    return {g, dg};
                }
                if (I.kind == 0) {
                    float g, dg;
                    moffat_generic_f32(C, k, &g, &dg);
                    return {g, dg};
                }
            }
            LeafD<double> m = moffat_bessel_d(I.kind, rec->p[0] - 1., double(k), double(ksq),
                                              rec->p[1]);
            return {T(m.phi), T(m.dphi)};
          }
        }
      }
      case GSMI_LEAF_KOLMOGOROV: {
        if (ksq == T(0)) return {T(1), T(0)};
        if constexpr (sizeof(T) == 4) {
            const float t = __builtin_amdgcn_exp2f((5.f / 6.f) * __builtin_amdgcn_logf(ksq));
            const float phi = __builtin_amdgcn_exp2f(-1.44269504088896341f * t);
            return {phi, ddiv(-5.f / 6.f * t, ksq) * phi};
        }
        T t = dexp(T(5. / 6.) * dlog(ksq));
        T phi = dexp(-t);
        return {phi, ddiv(T(-5. / 6.) * t, ksq) * phi};
      }
      case GSMI_LEAF_SERSIC: {
        if (ksq >= C.p[1]) {
            T irk;
            if constexpr (sizeof(T) == 4) irk = __builtin_amdgcn_rsqf(ksq);
            else irk = ddiv(T(1), dsqrt(ksq));
            T phi = ddiv(C.p[2] + C.p[3] * irk, ksq);
            return {phi, -ddiv(C.p[2] + T(1.5) * C.p[3] * irk, ksq * ksq)};
        }
        if (ksq < C.p[0])
            return {T(1) + ksq * (C.p[4] + ksq * C.p[5]), C.p[4] + T(2) * ksq * C.p[5]};
        T lk, dS, S;
        if constexpr (sizeof(T) == 4) {
            lk = 0.34657359027997264f * __builtin_amdgcn_logf(ksq);     // ln(ksq)/2
            S = spline_eq32_d(sersic_table<float>(tb, I.aux), I.ntab, lk, C.t0, C.t1, C.t2, &dS);
        } else {
            lk = T(0.5) * dlog(ksq);
            S = spline_eq<T>(sersic_table<T>(tb, I.aux), I.ntab, lk, &dS);
        }
        return {ddiv(S, ksq), ddiv(T(0.5) * dS - S, ksq * ksq)};
      }
      case GSMI_LEAF_BOX: {
        const T Xb = kxp * C.s;
        const T Yb = kyp * C.s2;
        T sx = fabs(Xb) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * Xb * Xb
                                  : ddiv(dsinpi(Xb), T(M_PI) * Xb);
        T sy = fabs(Yb) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * Yb * Yb
                                  : ddiv(dsinpi(Yb), T(M_PI) * Yb);
        return {sx * sy, T(0)};
      }
    }
    return {T(0), T(0)};
}

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

// ---------------------------------------------------------------------------------------------
// Kernel A: row DFTs of the residual.  grid (ceil(ny/rpg), nstamps); LDS 2 rpg N/2 complex.
// H(c) = sum_q R(q) e^{+2 pi i c q / N}, c = 0..N/2, from Z = DFT_{N/2}(R(2m) + i R(2m+1)):
// H(c) = E(c) + e^{2 pi i c/N} O(c), E = (Z(c) + conj Z(N/2-c)) / 2,
// O = (Z(c) - conj Z(N/2-c)) / 2i.
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
row_r2c(const gsmi_stamp* __restrict__ stamps, const T* __restrict__ resid,
        const cplx<T>* __restrict__ tw, cplx<T>* __restrict__ scratch)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    const gsmi_stamp* st = stamps + blockIdx.y;
    const int N = st->N;
    const int n2 = N / 2;
    const int W = n2 + 1;
    const int rpg = rows_per_group(N);
    const int r0 = blockIdx.x * rpg;
    const int nrow = min(rpg, st->ny - r0);
    if (nrow <= 0) return;
    const int nx = st->nx;
    const int x0 = st->x0;
    const T* in = resid + st->out_offset + size_t(r0) * nx;
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + rpg * n2;
    for (int e = threadIdx.x; e < nrow * n2; e += blockDim.x) bufa[e] = {T(0), T(0)};
    __syncthreads();
    T* za = reinterpret_cast<T*>(bufa);
    const float inv_nx = 1.f / float(nx);
    const float inv_n = 1.f / float(N);
    for (int e = threadIdx.x; e < nrow * nx; e += blockDim.x) {
        const int rl = fdiv(e, inv_nx);
        const int col = e - rl * nx;
        const int pq = x0 + col + N;                  // >= 0: x0 >= -N/2
        const int q = pq - N * fdiv(pq, inv_n);
        za[2 * (rl * n2 + (q >> 1)) + (q & 1)] = in[size_t(rl) * nx + col];
    }
    __syncthreads();
    const cplx<T>* res = fft_lds_c<T>(bufa, bufb, n2, nrow, tw, 2);
    cplx<T>* out = scratch + st->scratch_offset + size_t(r0) * W;
    const float inv_w = 1.f / float(W);
    for (int e = threadIdx.x; e < nrow * W; e += blockDim.x) {
        const int rl = fdiv(e, inv_w);
        const int c = e - rl * W;
        const cplx<T> zc = res[rl * n2 + (c == n2 ? 0 : c)];
        const cplx<T> zm = res[rl * n2 + (c == 0 ? 0 : n2 - c)];
        const cplx<T> E = {T(0.5) * (zc.x + zm.x), T(0.5) * (zc.y - zm.y)};
        const cplx<T> d = {zc.x - zm.x, zc.y + zm.y};              // Z(c) - conj Z(N/2-c)
        const cplx<T> O = {T(0.5) * d.y, T(-0.5) * d.x};          // d / 2i
        out[size_t(rl) * W + c] = cadd(E, cmul(tw[c], O));
    }
}

// ---------------------------------------------------------------------------------------------
// Kernel B: column DFTs of the row coefficients -> Rhat, then the gradient over every aliased
// k-image sample of the group's columns.  grid (ceil((N/2+1)/cpg), nstamps);
// LDS 2 cpg N complex.  ML = leaf slots with gradient accumulators (2 or 4).
// ---------------------------------------------------------------------------------------------
template <typename T, int ML, bool CPLX, bool SHAPE, int SIG>
__global__ void __launch_bounds__(kThreads)
grad_colfft(const gsmi_stamp* __restrict__ stamps, const gsmi_op* __restrict__ ops,
            const gsmi_leaf* __restrict__ leaves, Tables<T> tb, const cplx<T>* __restrict__ tw,
            const cplx<T>* __restrict__ scratch, double* __restrict__ partials)
{
    constexpr int PL = GSMI_GRAD_PER_LEAF;
    constexpr int NF = PL * ML + 3;
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    const gsmi_stamp* st = stamps + blockIdx.y;
    const int N = st->N;
    const int Nk = st->Nk;
    const int half = N / 2;
    const int W = half + 1;
    const int cpg = cols_per_group(N);
    const int c0 = blockIdx.x * cpg;
    const int ncol = max(0, min(cpg, W - c0));
    double* part = partials + (size_t(blockIdx.y) * gridDim.x + blockIdx.x) * GSMI_GRAD_FIELDS;
    const T dk = T(st->dk);
    const gsmi_op* sops = ops + st->op_begin;
    const int nops = st->op_count;
    const gsmi_leaf* sleaves = leaves + st->leaf_begin;
    const int hk = Nk / 2;
    const cplx<T>* X = scratch + st->axis_offset;
    const cplx<T>* Y = X + hk + 1 + hk;
    const cplx<T>* H = scratch + st->scratch_offset;
    const int ny = st->ny;
    const int y0 = st->y0;
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + cpg * N;

    for (int e = threadIdx.x; e < ncol * N; e += blockDim.x) bufa[e] = {T(0), T(0)};
    __syncthreads();
    const float inv_ncol = 1.f / float(max(ncol, 1));
    const float inv_n = 1.f / float(N);
    for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
        const int row = fdiv(e, inv_ncol);
        const int cl = e - row * ncol;
        const int py = y0 + row + N;                  // >= 0: y0 >= -N/2
        const int y = py - N * fdiv(py, inv_n);
        bufa[cl * N + y] = H[size_t(row) * W + c0 + cl];
    }
    __syncthreads();
    const cplx<T>* res = fft_lds_c<T>(bufa, bufb, N, ncol, tw, 1);

    // the program decoded into the LDS buffer the transform result does not occupy, and the
    // stamp's axis factors after it when they fit
    ProgramD d;
    cplx<T>* freeb = const_cast<cplx<T>*>(res == bufa ? bufb : bufa);
    LeafC<T>* lc = reinterpret_cast<LeafC<T>*>(freeb);
    bool ok = decode_program<T>(sops, nops, sleaves, tb, d, lc) && d.nl <= ML &&
              (CPLX || !d.has_phase);
    if constexpr (SIG > 0) ok = ok && matches_static<SIG>(sops, nops, sleaves);
    const int nl = (SIG > 0) ? kStaticPrograms[SIG].nl : d.nl;
    constexpr int kLeafSlots = kProgLeaves * sizeof(LeafC<T>) / sizeof(cplx<T>);
    const int naxis = 3 * hk + 2;
    const bool staged = kLeafSlots + naxis <= cpg * N;
    if (staged)
        for (int e = threadIdx.x; e < naxis; e += blockDim.x) freeb[kLeafSlots + e] = X[e];
    __syncthreads();
    // a specialised program's leaf constants in scalar registers; the interpreter reads LDS
    LeafC<T> lr[(SIG > 0) ? ML : 1];
    if constexpr (SIG > 0) {
#pragma unroll
        for (int l = 0; l < ML; ++l)
            if (l < nl) lr[l] = uniform_leaf(lc[l]);
    }
    auto leafc = [&](int l) -> const LeafC<T>& {
        if constexpr (SIG > 0) return lr[l];
        else return lc[l];
    };

    T acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = T(0);
    const T inv_n2 = T(1) / (T(N) * T(N));
    // the sample loop, with the axis factors in LDS or in global memory
    auto samples = [&](const cplx<T>* Xs, const cplx<T>* Ys) {
        for (int e = threadIdx.x; ok && e < ncol * N; e += blockDim.x) {
            const int cl = fdiv(e, inv_n);
            const int r = e - cl * N;
            const int c = c0 + cl;
            const T wc = (c == 0 || c == half) ? inv_n2 : T(2) * inv_n2;
            const cplx<T> A = {res[cl * N + r].x * wc, res[cl * N + r].y * wc};
            const int kx0 = c - N * fdiv(c + hk, inv_n);
            const int ky0 = r - N * fdiv(r + hk, inv_n);
            for (int kxa = kx0; kxa <= hk; kxa += N) {
                for (int kya = ky0; kya <= hk; kya += N) {
                    const bool neg = kxa < 0;
                    const int i = neg ? -kxa : kxa;
                    const int j = neg ? -kya : kya;
                    const T kx = T(i) * dk;
                    const T ky = T(j) * dk;
                    const cplx<T> seed = neg ? cconj(A) : A;
                    const cplx<T> Z = cmul(seed, cmul(Xs[i], Ys[j]));
                    T v[kProgLeaves], phi[ML], dv[ML];
    #pragma unroll
                    for (int l = 0; l < kProgLeaves; ++l) v[l] = T(0);
    #pragma unroll
                    for (int l = 0; l < ML; ++l) {
                        phi[l] = dv[l] = T(0);
                        if (l < nl) {
                            LeafI I = d.li[l];
                            if constexpr (SIG > 0) I.type = kStaticPrograms[SIG].type[l];
                            LeafD<T> q = eval_leafc_d<T>(leafc(l), I, sleaves + l, tb, kx, ky,
                                                         i, j, Nk);
                            phi[l] = q.phi;
                            v[l] = leafc(l).flux * q.phi;
                            dv[l] = leafc(l).flux * q.dphi;
                        }
                    }
                    cplx<T> FZ;
                    if constexpr (SIG > 0) {
                        const T P = static_ops<T, SIG>(v, -1, T(0));
                        FZ = {P * Z.x, P * Z.y};
                    } else if constexpr (CPLX) {
                        FZ = cmul(run_ops<T>(d, sops, v, -1, T(0), kx, ky), Z);
                    } else {
                        const T P = run_ops_real<T>(d, sops, v, -1, T(0));
                        FZ = {P * Z.x, P * Z.y};
                    }
                    acc[PL * ML] += kx * FZ.y;
                    acc[PL * ML + 1] += ky * FZ.y;
                    acc[PL * ML + 2] += FZ.x;
    #pragma unroll
                    for (int l = 0; l < ML; ++l) {
                        if (l < nl) {
                            T tx;          // Re[(dP/dv_l) Z]
                            if constexpr (SIG > 0) {
                                tx = (static_ops<T, SIG>(v, l, T(1)) -
                                      static_ops<T, SIG>(v, l, T(0))) * Z.x;
                            } else if constexpr (CPLX) {
                                const cplx<T> p1 = run_ops<T>(d, sops, v, l, T(1), kx, ky);
                                const cplx<T> p0 = run_ops<T>(d, sops, v, l, T(0), kx, ky);
                                tx = cmul(csub(p1, p0), Z).x;
                            } else {
                                tx = (run_ops_real<T>(d, sops, v, l, T(1)) -
                                      run_ops_real<T>(d, sops, v, l, T(0))) * Z.x;
                            }
                            const LeafC<T>& L = leafc(l);
                            T kxp, kyp;
                            leaf_kprime(L, d.li[l], kx, ky, &kxp, &kyp);
                            const T sc = L.s;
                            const T g = tx * dv[l];
                            const T g2 = T(2) * sc * sc * g;
                            acc[PL * l] += tx * phi[l];
                            acc[PL * l + 1] += T(2) * sc * (kxp * kxp + kyp * kyp) * g;
                            acc[PL * l + 2] += g2 * kxp * kx;
                            acc[PL * l + 3] += g2 * kyp * kx;
                            acc[PL * l + 4] += g2 * kxp * ky;
                            acc[PL * l + 5] += g2 * kyp * ky;
                            if constexpr (SHAPE) {
                                const int type = (SIG > 0) ? kStaticPrograms[SIG].type[l]
                                                           : d.li[l].type;
                                if (type == GSMI_LEAF_MOFFAT) {
                                    const T kk = sc * dsqrt(kxp * kxp + kyp * kyp);
                                    T dg;
                                    if constexpr (std::is_same<T, float>::value)
                                        dg = moffat_dg_dnu_f32(float(sleaves[l].p[0]) - 1.f, kk);
                                    else
                                        dg = moffat_dg_dnu(sleaves[l].p[0] - 1., kk);
                                    acc[PL * l + 6] += tx * L.flux * dg;
                                }
                            }
                        }
                    }
                }
            }
        }
    };
    if (staged)
        samples(freeb + kLeafSlots, freeb + kLeafSlots + 2 * hk + 1);
    else
        samples(X, Y);
    __syncthreads();                       // LDS is reused for the reduction
    double* red = reinterpret_cast<double*>(smem_raw);
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int nwave = blockDim.x / 64;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        double x = wave_sum(double(acc[f]));
        if (lane == 0) red[wave * NF + f] = x;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < GSMI_GRAD_FIELDS; f += blockDim.x) {
        // output layout: leaf slots at PL l for l < GSMI_GRAD_MAXLEAF, then the 3 whole-result
        // fields; slots beyond ML read zero
        int src = -1;
        if (f < PL * ML) src = f;
        else if (f >= PL * GSMI_GRAD_MAXLEAF) src = PL * ML + (f - PL * GSMI_GRAD_MAXLEAF);
        double x = 0.;
        if (!ok) x = NAN;
        else if (src >= 0)
            for (int w = 0; w < nwave; ++w) x += red[w * NF + src];
        part[f] = x;
    }
}

// Kernel C: grad[s][f] = sum over the stamp's workgroups of partials, in order.
__global__ void __launch_bounds__(kThreads)
grad_reduce(const double* __restrict__ partials, int ngroups, int nstamps,
            double* __restrict__ grad)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nstamps * GSMI_GRAD_FIELDS) return;
    const int s = idx / GSMI_GRAD_FIELDS;
    const int f = idx - s * GSMI_GRAD_FIELDS;
    double x = 0.;
    for (int g = 0; g < ngroups; ++g)
        x += partials[(size_t(s) * ngroups + g) * GSMI_GRAD_FIELDS + f];
    grad[idx] = x;
}

int adjoint_groups(int N) { return (N / 2 + 1 + cols_per_group(N) - 1) / cols_per_group(N); }

template <typename T, bool CPLX, bool SHAPE>
void launch_grad(dim3 gb, size_t ldsb, hipStream_t s, int max_leaves, const gsmi_stamp* stamps,
                 const gsmi_op* ops, const gsmi_leaf* leaves, Tables<T> tb,
                 const cplx<T>* tw, const void* scratch, double* partials)
{
    const cplx<T>* sc = reinterpret_cast<const cplx<T>*>(scratch);
    if (max_leaves <= 2)
        hipLaunchKernelGGL((grad_colfft<T, 2, CPLX, SHAPE, 0>), gb, dim3(kThreads), ldsb, s,
                           stamps, ops, leaves, tb, tw, sc, partials);
    else
        hipLaunchKernelGGL((grad_colfft<T, 4, CPLX, SHAPE, 0>), gb, dim3(kThreads), ldsb, s,
                           stamps, ops, leaves, tb, tw, sc, partials);
}

// The specialised adjoint of program shape sig (phase-free by construction); false if there is
// no such kernel and the interpreter should run.
template <typename T>
bool launch_grad_static(int sig, bool shape, dim3 gb, size_t ldsb, hipStream_t s,
                        const gsmi_stamp* stamps, const gsmi_op* ops, const gsmi_leaf* leaves,
                        Tables<T> tb, const cplx<T>* tw, const void* scratch, double* partials)
{
    const cplx<T>* sc = reinterpret_cast<const cplx<T>*>(scratch);
    switch (sig) {
      case 1:
        hipLaunchKernelGGL((grad_colfft<T, 4, false, false, 1>), gb, dim3(kThreads), ldsb, s,
                           stamps, ops, leaves, tb, tw, sc, partials);
        return true;
      case 2:
        if (shape)
            hipLaunchKernelGGL((grad_colfft<T, 2, false, true, 2>), gb, dim3(kThreads), ldsb, s,
                               stamps, ops, leaves, tb, tw, sc, partials);
        else
            hipLaunchKernelGGL((grad_colfft<T, 2, false, false, 2>), gb, dim3(kThreads), ldsb,
                               s, stamps, ops, leaves, tb, tw, sc, partials);
        return true;
      case 3:
        hipLaunchKernelGGL((grad_colfft<T, 2, false, false, 3>), gb, dim3(kThreads), ldsb, s,
                           stamps, ops, leaves, tb, tw, sc, partials);
        return true;
      case 4:
        hipLaunchKernelGGL((grad_colfft<T, 2, false, false, 4>), gb, dim3(kThreads), ldsb, s,
                           stamps, ops, leaves, tb, tw, sc, partials);
        return true;
    }
    return false;
}

template <typename T>
int adjoint_impl(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, int max_leaves,
                 int flags, const gsmi_op* ops, const gsmi_leaf* leaves, Tables<T> tb,
                 const T* resid,
                 void* scratch, double* partials, double* grad, void* stream)
{
    if (!fft_size_ok(N, sizeof(cplx<T>))) {
        set_error("unsupported FFT size " + std::to_string(N) +
                  " (need 2^k or 3*2^k, <= 2048 in FP64, <= 4096 in FP32)");
        return GSMI_E_SIZE;
    }
    if (max_leaves < 1 || max_leaves > GSMI_GRAD_MAXLEAF) {
        set_error("adjoint: the program must have 1.." + std::to_string(GSMI_GRAD_MAXLEAF) +
                  " leaves");
        return GSMI_E_ARG;
    }
    if (nstamps <= 0) return GSMI_OK;
    const cplx<T>* tw = twiddles<T>(N);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc = launch_axis_factors<T>(stamps, nstamps, ops, leaves, scratch, stream);
    if (rc) return rc;
    const int rpg = rows_per_group(N);
    dim3 ga((max_ny + rpg - 1) / rpg, nstamps);
    size_t ldsa = 2 * size_t(rpg) * (N / 2) * sizeof(cplx<T>);
    hipLaunchKernelGGL(row_r2c<T>, ga, dim3(kThreads), ldsa, s, stamps, resid, tw,
                       reinterpret_cast<cplx<T>*>(scratch));
    rc = hip_check(hipGetLastError(), "row_r2c launch");
    if (rc) return rc;
    const int cpg = cols_per_group(N);
    const int ng = adjoint_groups(N);
    dim3 gb(ng, nstamps);
    size_t ldsb = 2 * size_t(cpg) * N * sizeof(cplx<T>);
    const bool cx = flags & GSMI_ADJ_COMPLEX;
    const bool sh = flags & GSMI_ADJ_SHAPE;
    const int sig = (flags >> 8) & 0xff;
    if (!cx && sig > 0 &&
        launch_grad_static<T>(sig, sh, gb, ldsb, s, stamps, ops, leaves, tb, tw, scratch,
                              partials)) {
        // specialised kernel launched
    } else if (cx && sh)
        launch_grad<T, true, true>(gb, ldsb, s, max_leaves, stamps, ops, leaves, tb, tw,
                                   scratch, partials);
    else if (cx)
        launch_grad<T, true, false>(gb, ldsb, s, max_leaves, stamps, ops, leaves, tb, tw,
                                    scratch, partials);
    else if (sh)
        launch_grad<T, false, true>(gb, ldsb, s, max_leaves, stamps, ops, leaves, tb, tw,
                                    scratch, partials);
    else
        launch_grad<T, false, false>(gb, ldsb, s, max_leaves, stamps, ops, leaves, tb, tw,
                                     scratch, partials);
    rc = hip_check(hipGetLastError(), "grad_colfft launch");
    if (rc) return rc;
    const int total = nstamps * GSMI_GRAD_FIELDS;
    hipLaunchKernelGGL(grad_reduce, dim3((total + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                       partials, ng, nstamps, grad);
    return hip_check(hipGetLastError(), "grad_reduce launch");
}

}  // namespace
}  // namespace gsmi

using namespace gsmi;

extern "C" {

int gsmi_adjoint_partials(int N, int nstamps, int64_t* ndoubles)
{
    if (!ndoubles || nstamps < 0 || !fft_size_ok(N, sizeof(cplx<float>))) {
        set_error("gsmi_adjoint_partials: bad size");
        return GSMI_E_ARG;
    }
    *ndoubles = int64_t(adjoint_groups(N)) * nstamps * GSMI_GRAD_FIELDS;
    return GSMI_OK;
}

int gsmi_adjoint_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, int max_leaves,
                     int flags, const gsmi_op* ops, const gsmi_leaf* leaves,
                     const double* tables, const int32_t* cliprows, const double* resid,
                     void* scratch, double* partials, double* grad, void* stream)
{
    if (!stamps || !ops || !leaves || !resid || !scratch || !partials || !grad || max_ny <= 0) {
        set_error("gsmi_adjoint_f64: null buffer or empty stamp");
        return GSMI_E_ARG;
    }
    Tables<double> tb{tables, cliprows};
    return adjoint_impl<double>(stamps, nstamps, N, max_ny, max_leaves, flags, ops,
                                leaves, tb, resid,
                                scratch, partials, grad, stream);
}

int gsmi_adjoint_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, int max_leaves,
                     int flags, const gsmi_op* ops, const gsmi_leaf* leaves,
                     const double* tables, const float* tables32, const float* resid,
                     void* scratch, double* partials, double* grad, void* stream)
{
    if (!stamps || !ops || !leaves || !resid || !scratch || !partials || !grad || max_ny <= 0) {
        set_error("gsmi_adjoint_f32: null buffer or empty stamp");
        return GSMI_E_ARG;
    }
    Tables<float> tb{tables, tables32, reinterpret_cast<const float4*>(tables32)};
    return adjoint_impl<float>(stamps, nstamps, N, max_ny, max_leaves, flags, ops,
                               leaves, tb, resid,
                               scratch, partials, grad, stream);
}

}
