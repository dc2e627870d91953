// Device-side program evaluation shared by the forward and adjoint kernels: the leaf k-space
// formulas, the register-stack program of include/gsmi.h, the hoisted axis factors, and the
// in-LDS Stockham FFT.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/gsmi.h"
#include "device_math.h"

namespace gsmi {

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------------------------
// Leaf evaluation
// ---------------------------------------------------------------------------------------------
template <typename T>
struct Tables;

template <>
struct Tables<double> {
    const double* tab;
    const int32_t* clip;
};
template <>
struct Tables<float> {
    const double* tab;
    const float* tab32;
};

template <typename T>
__device__ inline T sersic_spline(const gsmi_leaf& L, const Tables<T>& tb, T lk);

template <>
__device__ inline double sersic_spline<double>(const gsmi_leaf& L, const Tables<double>& tb,
                                               double lk)
{
    // Table::lookup on equally spaced log-k knots (src/Table.cpp:77-92, 560-577)
    const int n = L.ntab;
    const double* args = tb.tab + L.table;
    const double* vals = args + n;
    const double* y2 = vals + n;
    double a0 = args[0];
    double da = (args[n - 1] - a0) / (n - 1);
    if (lk < a0 || lk > args[n - 1]) return 0.;
    int i = int(ceil((lk - a0) / da));
    if (i >= n) i = n - 1;
    if (i <= 0) i = 1;
    double h = args[i] - args[i - 1];
    double aa = args[i] - lk;
    double bb = h - aa;
    return (aa * vals[i - 1] + bb * vals[i] -
            (1. / 6.) * aa * bb * ((aa + h) * y2[i - 1] + (bb + h) * y2[i])) / h;
}

template <>
__device__ inline float sersic_spline<float>(const gsmi_leaf& L, const Tables<float>& tb, float lk)
{
    const int n = L.ntab;
    const float* args = tb.tab32 + L.table;
    const float* vals = args + n;
    const float* y2 = vals + n;
    float a0 = args[0];
    float inv_da = float(n - 1) / (args[n - 1] - a0);
    if (lk < a0 || lk > args[n - 1]) return 0.f;
    int i = int(ceilf((lk - a0) * inv_da));
    i = min(max(i, 1), n - 1);
    float h = args[i] - args[i - 1];
    float aa = args[i] - lk;
    float bb = h - aa;
    return (aa * vals[i - 1] + bb * vals[i] -
            (1.f / 6.f) * aa * bb * ((aa + h) * y2[i - 1] + (bb + h) * y2[i])) / h;
}

template <typename T>
__device__ inline bool exp_clipped(const gsmi_leaf& L, const Tables<T>& tb, int i, int j, int Nk,
                                   T ksq);

template <>
__device__ inline bool exp_clipped<double>(const gsmi_leaf& L, const Tables<double>& tb, int i,
                                           int j, int Nk, double ksq)
{
    if (L.cliprows >= 0) {
        // Row ranges computed on the host exactly as GetKValueRange1d/2d
        // (src/SBProfile.cpp:508-588) for the stored k-image rows.
        const int32_t* r = tb.clip + L.cliprows + 2 * (j + Nk / 2);
        return i < r[0] || i >= r[1];
    }
    return ksq > L.p[0];
}

template <>
__device__ inline bool exp_clipped<float>(const gsmi_leaf& L, const Tables<float>&, int, int, int,
                                          float ksq)
{
    return ksq > float(L.p[0]);
}

// Moffat kV for the Bessel-function cases, kept out of line so the common path stays small.
__device__ __noinline__ double moffat_kv_bessel(int kind, double nu, double k, double ksq,
                                                double at_zero)
{
    switch (kind) {
      case 2: return ksq == 0. ? 1. : bessel_k(1., k) * k;
      case 3: return ksq == 0. ? 2. : bessel_k(2., k) * ksq;
      case 4: return ksq == 0. ? 8. : bessel_k(3., k) * k * ksq;
      default: return ksq == 0. ? at_zero : bessel_k(nu, k) * pow(k, nu);
    }
}

// Value of leaf L at k-image sample (i, j) (kx = i dk, ky = j dk).
template <typename T>
__device__ inline T eval_leaf(const gsmi_leaf* __restrict__ Lp, const Tables<T>& tb, T kx, T ky,
                              int i, int j, int Nk)
{
    const gsmi_leaf& L = *Lp;
    T kxp, kyp;
    if (L.separable) {
        kxp = T(L.jac[0]) * kx;
        kyp = T(L.jac[3]) * ky;
    } else {
        kxp = T(L.jac[0]) * kx + T(L.jac[2]) * ky;
        kyp = T(L.jac[1]) * kx + T(L.jac[3]) * ky;
    }
    const T s = T(L.scale);
    const T X = kxp * s;
    const T Y = kyp * s;
    const T flux = T(L.flux);
    switch (L.type) {
      case GSMI_LEAF_EXPONENTIAL: {
        // SBExponential::fillKImage (src/SBExponential.cpp:442-510)
        T ksq = X * X + Y * Y;
        if (exp_clipped<T>(L, tb, i, j, Nk, ksq)) return T(0);
        T q = T(1) + ksq;
        return flux / (q * dsqrt(q));
      }
      case GSMI_LEAF_GAUSSIAN: {
        // SBGaussian::fillKImage (src/SBGaussian.cpp:200-290)
        if (L.separable) return flux * dexp(T(-0.5) * X * X) * dexp(T(-0.5) * Y * Y);
        T ksq = X * X + Y * Y;
        if (ksq > T(L.p[0])) return T(0);
        if (ksq < T(L.p[1])) return flux * (T(1) - T(0.5) * ksq * (T(1) - T(0.25) * ksq));
        return flux * dexp(T(-0.5) * ksq);
      }
      case GSMI_LEAF_MOFFAT: {
        // SBMoffat kV_* (src/SBMoffat.cpp:286-345); flux holds knorm
        T ksq = X * X + Y * Y;
        T k = dsqrt(ksq);
        switch (L.kind) {
          case 15: return flux * dexp(-k);
          case 25: return flux * dexp(-k) * (T(1) + k);
          case 35: return flux * dexp(-k) * (T(3) + (T(3) + k) * k);
          default:
            if constexpr (sizeof(T) == 4) {
                if (L.kind == 2 || L.kind == 3 || L.kind == 4) {
                    float g, dg;
                    moffat_int_kind_f32(L.kind, k, &g, &dg);
                    return flux * g;
                }
            }
            return flux * T(moffat_kv_bessel(L.kind, L.p[0] - 1., double(k), double(ksq), L.p[1]));
        }
      }
      case GSMI_LEAF_KOLMOGOROV: {
        // SBKolmogorov::fillKImage, KolmogorovInfo::kValue (src/SBKolmogorov.cpp:185-259)
        T ksq = X * X + Y * Y;
        return flux * dexp(-dexp(T(5. / 6.) * dlog(ksq)));
      }
      case GSMI_LEAF_SERSIC: {
        // SersicInfo::kValue (src/SBSersic.cpp:363-376)
        T ksq = X * X + Y * Y;
        if (ksq >= T(L.p[1])) return flux * (T(L.p[2]) + T(L.p[3]) / dsqrt(ksq)) / ksq;
        if (ksq < T(L.p[0])) return flux * (T(1) + ksq * (T(L.p[4]) + ksq * T(L.p[5])));
        T lk = T(0.5) * dlog(ksq);
        return flux * sersic_spline<T>(L, tb, lk) / ksq;
      }
      case GSMI_LEAF_BOX: {
        // SBBox::fillKImage with math::sinc (src/SBBox.cpp:151-224, src/math/Sinc.cpp:29-33)
        const T Xb = kxp * T(L.scale);
        const T Yb = kyp * T(L.scale2);
        T sx = fabs(Xb) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * Xb * Xb
                                  : dsinpi(Xb) / (T(M_PI) * Xb);
        T sy = fabs(Yb) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * Yb * Yb
                                  : dsinpi(Yb) / (T(M_PI) * Yb);
        return flux * sx * sy;
      }
    }
    return T(0);
}

// A small stack held in registers: push/pop shift every slot, so all indices are static.
template <typename T, int S>
struct RegStack {
    cplx<T> s[S];
    __device__ __forceinline__ void push(cplx<T> v)
    {
#pragma unroll
        for (int k = S - 1; k > 0; --k) s[k] = s[k - 1];
        s[0] = v;
    }
    __device__ __forceinline__ cplx<T> pop()
    {
        cplx<T> v = s[0];
#pragma unroll
        for (int k = 0; k < S - 1; ++k) s[k] = s[k + 1];
        return v;
    }
};

template <typename T>
struct StackDepth;
template <>
struct StackDepth<double> { static constexpr int value = GSMI_STACK_F64; };
template <>
struct StackDepth<float> { static constexpr int value = GSMI_STACK_F32; };

// Evaluate the stamp program at k-image sample (i, j).
template <typename T>
__device__ inline cplx<T> eval_program(const gsmi_op* __restrict__ ops, int nops,
                                       const gsmi_leaf* __restrict__ leaves, const Tables<T>& tb,
                                       T dk, int i, int j, int Nk)
{
    RegStack<T, StackDepth<T>::value> st;
    const T kx = T(i) * dk;
    const T ky = T(j) * dk;
    for (int o = 0; o < nops; ++o) {
        const gsmi_op* op = ops + o;
        switch (op->code) {
          case GSMI_OP_LEAF: {
            T v = eval_leaf<T>(leaves + op->arg, tb, kx, ky, i, j, Nk);
            st.push({v, T(0)});
            break;
          }
          case GSMI_OP_MUL: {
            cplx<T> b = st.pop();
            st.s[0] = cmul(st.s[0], b);
            break;
          }
          case GSMI_OP_ADD: {
            cplx<T> b = st.pop();
            st.s[0] = cadd(st.s[0], b);
            break;
          }
          case GSMI_OP_SCALE:
            st.s[0].x *= T(op->a);
            st.s[0].y *= T(op->a);
            break;
          case GSMI_OP_PHASE: {
            T arg = T(op->a) * kx + T(op->b) * ky;
            T sn, cs;
            dsincos(arg, &sn, &cs);
            const T fs = T(op->c);
            st.s[0] = cmul(st.s[0], cplx<T>{fs * cs, -fs * sn});
            break;
          }
        }
    }
    return st.s[0];
}

// Separable factors hoisted out of the program (pixel sinc with a diagonal Jacobian, the phase
// of a shift, constant scalings): X(kx) for the stored kx = 0..Nk/2 and Y(ky) for ky =
// -Nk/2..Nk/2.  Constants and flux go into X.
template <typename T>
__device__ inline cplx<T> axis_factor(const gsmi_op* __restrict__ ops, int nops,
                                      const gsmi_leaf* __restrict__ leaves, T k, bool xaxis)
{
    cplx<T> f = {T(1), T(0)};
    for (int o = 0; o < nops; ++o) {
        const gsmi_op* op = ops + o;
        switch (op->code) {
          case GSMI_OP_LEAF: {
            const gsmi_leaf* L = leaves + op->arg;     // a Box with a diagonal Jacobian
            const T u = xaxis ? T(L->jac[0]) * k * T(L->scale) : T(L->jac[3]) * k * T(L->scale2);
            T sc = fabs(u) < T(1e-4) ? T(1) - T(M_PI * M_PI / 6.) * u * u
                                     : dsinpi(u) / (T(M_PI) * u);
            if (xaxis) sc *= T(L->flux);
            f.x *= sc;
            f.y *= sc;
            break;
          }
          case GSMI_OP_SCALE:
            if (xaxis) { f.x *= T(op->a); f.y *= T(op->a); }
            break;
          case GSMI_OP_PHASE: {
            T arg = (xaxis ? T(op->a) : T(op->b)) * k;
            T sn, cs;
            dsincos(arg, &sn, &cs);
            const T fs = xaxis ? T(op->c) : T(1);
            f = cmul(f, cplx<T>{fs * cs, -fs * sn});
            break;
          }
        }
    }
    return f;
}

// ---------------------------------------------------------------------------------------------
// In-LDS Stockham FFT (inverse, e^{+2 pi i}) over nseg contiguous segments of length n.
// tw[m * twstride] = exp(+2 pi i m / n).  Mixed radix 4/2/3; returns the buffer with the result.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ inline void bfly2(cplx<T>* v)
{
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T>
__device__ inline void bfly4(cplx<T>* v)
{
    cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cplx<T> a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    cplx<T> a3 = {-d.y, d.x};   // i * (x1 - x3)
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <typename T>
__device__ inline void bfly3(cplx<T>* v)
{
    const T h = T(0.86602540378443864676);   // sqrt(3)/2
    cplx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cplx<T> m = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
    cplx<T> id = {-h * d.y, h * d.x};        // i sqrt(3)/2 (x1 - x2)
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, id);
    v[2] = csub(m, id);
}

template <typename T, int R>
__device__ inline void stockham_pass(const cplx<T>* src, cplx<T>* dst, int n, int nseg, int Ns,
                                     const cplx<T>* tw, int twstride)
{
    const int nb = n / R;
    const int tstep = (n / (Ns * R)) * twstride;
    for (int t = threadIdx.x; t < nseg * nb; t += blockDim.x) {
        const int seg = t / nb;
        const int tt = t - seg * nb;
        const int jj = tt % Ns;
        const cplx<T>* s = src + seg * n + tt;
        cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[jj * r * tstep]);
        }
        if (R == 2) bfly2(v);
        if (R == 3) bfly3(v);
        if (R == 4) bfly4(v);
        cplx<T>* d = dst + seg * n + (tt / Ns) * Ns * R + jj;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * Ns] = v[r];
    }
}

template <typename T>
__device__ cplx<T>* fft_lds(cplx<T>* a, cplx<T>* b, int n, int nseg, const cplx<T>* tw,
                            int twstride)
{
    int Ns = 1;
    int rem = n;
    cplx<T>* src = a;
    cplx<T>* dst = b;
    while (rem > 1) {
        if (rem % 3 == 0) {
            stockham_pass<T, 3>(src, dst, n, nseg, Ns, tw, twstride);
            Ns *= 3; rem /= 3;
        } else if (rem % 4 == 0) {
            stockham_pass<T, 4>(src, dst, n, nseg, Ns, tw, twstride);
            Ns *= 4; rem /= 4;
        } else {
            stockham_pass<T, 2>(src, dst, n, nseg, Ns, tw, twstride);
            Ns *= 2; rem /= 2;
        }
        __syncthreads();
        cplx<T>* tmp = src; src = dst; dst = tmp;
    }
    return src;
}

__host__ __device__ inline int cols_per_group(int N)
{
    return N >= 1024 ? 1 : 1024 / N;
}
__host__ __device__ inline int rows_per_group(int N)
{
    int n2 = N / 2;
    return n2 >= 1024 ? 1 : 1024 / n2;
}

}  // namespace gsmi
