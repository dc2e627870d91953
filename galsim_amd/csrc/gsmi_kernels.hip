// HIP kernels for GalSim's FFT rendering path on gfx950, and the device half of the C ABI.
//
// Data flow for one batch of stamps that share an FFT size N (all launches on one stream):
//   1. fold_colfft: each workgroup evaluates the stamp's k-space program on a group of kx columns
//      of the wrapped N-grid -- every aliased k-image sample (kx, ky) congruent to the target cell
//      modulo N, with conj(F(-k)) for kx < 0 -- straight into LDS (the fused product / sinc /
//      phase / wrap pass), runs the inverse DFT along ky for those columns in LDS (Stockham
//      radix 4/2/3), and writes the rows the output stamp needs to a scratch buffer.
//   2. row_c2r: each workgroup takes rows of that scratch, forms the half-length complex
//      sequence of the real inverse transform, runs it in LDS, and writes the cropped, 1/N^2
//      scaled pixels.
// The k-image of the reference never exists in memory; nor does the full N x N real image.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsmi.h"
#include "device_math.h"

namespace gsmi {

void set_error(const std::string& msg);

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

// ---------------------------------------------------------------------------------------------
// Kernel 0: per-stamp separable axis factors X(kx), Y(ky) into the workspace (grid: nstamps)
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
axis_factors(const gsmi_stamp* __restrict__ stamps, const gsmi_op* __restrict__ ops,
             const gsmi_leaf* __restrict__ leaves, cplx<T>* __restrict__ ws)
{
    const gsmi_stamp* st = stamps + blockIdx.x;
    const int hk = st->Nk / 2;
    const T dk = T(st->dk);
    const gsmi_op* aops = ops + st->op_begin + st->op_count;
    const gsmi_leaf* sleaves = leaves + st->leaf_begin;
    cplx<T>* out = ws + st->axis_offset;
    const int total = (hk + 1) + (2 * hk + 1);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        if (e <= hk) out[e] = axis_factor<T>(aops, st->axis_count, sleaves, T(e) * dk, true);
        else out[e] = axis_factor<T>(aops, st->axis_count, sleaves, T(e - hk - 1 - hk) * dk, false);
    }
}

// ---------------------------------------------------------------------------------------------
// Kernel 1: fused k-space evaluation + product/sinc/phase + Hermitian wrap + column inverse DFT
// grid: (ceil((N/2+1)/cpg), nstamps); dynamic LDS: 2 * cpg * N complex
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
fold_colfft(const gsmi_stamp* __restrict__ stamps, const gsmi_op* __restrict__ ops,
            const gsmi_leaf* __restrict__ leaves, Tables<T> tb, const cplx<T>* __restrict__ tw,
            cplx<T>* __restrict__ scratch)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    const gsmi_stamp* st = stamps + blockIdx.y;
    const int N = st->N;
    const int Nk = st->Nk;
    const int half = N / 2;
    const int cpg = cols_per_group(N);
    const int c0 = blockIdx.x * cpg;
    const int ncol = min(cpg, half + 1 - c0);
    if (ncol <= 0) return;
    const T dk = T(st->dk);
    const gsmi_op* sops = ops + st->op_begin;
    const int nops = st->op_count;
    const gsmi_leaf* sleaves = leaves + st->leaf_begin;
    const int hk = Nk / 2;
    const cplx<T>* X = scratch + st->axis_offset;
    const cplx<T>* Y = X + hk + 1 + hk;               // Y[ky] for ky in [-hk, hk]
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + cpg * N;

    for (int e = threadIdx.x; e < ncol * N; e += blockDim.x) {
        const int cl = e / N;
        const int r = e - cl * N;
        const int c = c0 + cl;
        cplx<T> acc = {T(0), T(0)};
        const int kx0 = c - N * ((c + hk) / N);
        const int ky0 = r - N * ((r + hk) / N);
        for (int kx = kx0; kx <= hk; kx += N) {
            for (int ky = ky0; ky <= hk; ky += N) {
                // the stored sample (kx >= 0); its conjugate stands for (kx, ky) when kx < 0
                const bool neg = kx < 0;
                const int i = neg ? -kx : kx;
                const int j = neg ? -ky : ky;
                cplx<T> v = eval_program<T>(sops, nops, sleaves, tb, dk, i, j, Nk);
                v = cmul(v, cmul(X[i], Y[j]));
                acc.x += v.x;
                acc.y += neg ? -v.y : v.y;
            }
        }
        bufa[cl * N + r] = acc;
    }
    __syncthreads();
    cplx<T>* res = fft_lds<T>(bufa, bufb, N, ncol, tw, 1);
    // rows of the output window: position p <-> DFT index p mod N
    cplx<T>* out = scratch + st->scratch_offset;
    const int W = half + 1;
    const int ny = st->ny;
    const int y0 = st->y0;
    for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
        const int row = e / ncol;
        const int cl = e - row * ncol;
        const int p = y0 + row;
        const int y = (p % N + N) % N;
        out[row * W + c0 + cl] = res[cl * N + y];
    }
}

// ---------------------------------------------------------------------------------------------
// Kernel 2: real inverse DFT along x of the column-transformed rows, crop and scale.
// Row z holds Z(c), c = 0..N/2; the real sequence x(q) = sum_c Z(c) e^{2 pi i c q / N} (Hermitian
// extension) is formed as v = IDFT_{N/2}(E + i O), x(2m) = Re v(m), x(2m+1) = Im v(m), with
// E(c) = Z(c) + conj(Z(N/2-c)) and O(c) = (Z(c) - conj(Z(N/2-c))) e^{2 pi i c / N}.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ void row_c2r_body(cplx<T>* lds, int rpg, const cplx<T>* __restrict__ in, int nrow,
                             int N, int x0, int nx, const cplx<T>* __restrict__ tw,
                             T* __restrict__ out)
{
    const int n2 = N / 2;
    const int W = n2 + 1;
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + rpg * n2;
    for (int e = threadIdx.x; e < nrow * n2; e += blockDim.x) {
        const int rl = e / n2;
        const int c = e - rl * n2;
        const cplx<T>* z = in + size_t(rl) * W;
        cplx<T> zc = z[c];
        cplx<T> zm = z[n2 - c];
        if (c == 0) { zc.y = T(0); zm.y = T(0); }    // DC and Nyquist are real
        cplx<T> E = cadd(zc, cconj(zm));
        cplx<T> O = cmul(csub(zc, cconj(zm)), tw[c]);
        bufa[rl * n2 + c] = {E.x - O.y, E.y + O.x};  // E + i O
    }
    __syncthreads();
    cplx<T>* res = fft_lds<T>(bufa, bufb, n2, nrow, tw, 2);
    const T norm = T(1) / (T(N) * T(N));
    for (int e = threadIdx.x; e < nrow * nx; e += blockDim.x) {
        const int rl = e / nx;
        const int col = e - rl * nx;
        const int q = ((x0 + col) % N + N) % N;
        const cplx<T> v = res[rl * n2 + (q >> 1)];
        out[size_t(rl) * nx + col] = ((q & 1) ? v.y : v.x) * norm;
    }
}

// grid: (ceil(ny/rpg), nstamps); dynamic LDS: 2 * rpg * N/2 complex
template <typename T>
__global__ void __launch_bounds__(kThreads)
row_c2r(const gsmi_stamp* __restrict__ stamps, const cplx<T>* __restrict__ scratch,
        const cplx<T>* __restrict__ tw, T* __restrict__ out)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    const gsmi_stamp* st = stamps + blockIdx.y;
    const int N = st->N;
    const int rpg = rows_per_group(N);
    const int r0 = blockIdx.x * rpg;
    const int nrow = min(rpg, st->ny - r0);
    if (nrow <= 0) return;
    const int W = N / 2 + 1;
    row_c2r_body<T>(lds, rpg, scratch + st->scratch_offset + size_t(r0) * W, nrow, N, st->x0,
                    st->nx, tw, out + st->out_offset + size_t(r0) * st->nx);
}

// full N x N real image from N column-transformed rows (irfft entry point)
__global__ void __launch_bounds__(kThreads)
irfft_rows_f64(const cplx<double>* __restrict__ scratch, int N, const cplx<double>* __restrict__ tw,
               double* __restrict__ out)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    const int rpg = rows_per_group(N);
    const int r0 = blockIdx.x * rpg;
    const int nrow = min(rpg, N - r0);
    if (nrow <= 0) return;
    const int W = N / 2 + 1;
    row_c2r_body<double>(lds, rpg, scratch + size_t(r0) * W, nrow, N, -N / 2, N, tw,
                         out + size_t(r0) * N);
}

// ---------------------------------------------------------------------------------------------
// k-image evaluation (drawKImage / SBProfile::drawK): one thread per sample, coalesced stores
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
draw_kimage_f64(const gsmi_op* __restrict__ ops, int nops, const gsmi_leaf* __restrict__ leaves,
                Tables<double> tb, int xmin, int ymin, int nx, int ny, double dk, int Nk,
                cplx<double>* __restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny) return;
    const int row = idx / nx;
    const int col = idx - row * nx;
    out[idx] = eval_program<double>(ops, nops, leaves, tb, dk, xmin + col, ymin + row, Nk);
}

// wrapImage(hermx) of a stored (Nk/2+1) x (Nk+1) k-image onto (N/2+1) x N, rows ky=-N/2..N/2-1
__global__ void __launch_bounds__(kThreads)
wrap_hermx_f64(const cplx<double>* __restrict__ kin, int Nk, int N, cplx<double>* __restrict__ kout)
{
    const int W = N / 2 + 1;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= W * N) return;
    const int row = idx / W;
    const int c = idx - row * W;
    const int r = row - N / 2;                 // ky representative
    const int hk = Nk / 2;
    const int Wk = hk + 1;
    const int kx0 = c - N * ((c + hk) / N);
    const int ky0 = ((r % N) + N) % N;
    const int ky00 = ky0 - N * ((ky0 + hk) / N);
    cplx<double> acc = {0., 0.};
    for (int kx = kx0; kx <= hk; kx += N) {
        for (int ky = ky00; ky <= hk; ky += N) {
            if (kx >= 0) acc = cadd(acc, kin[(ky + hk) * Wk + kx]);
            else acc = cadd(acc, cconj(kin[(-ky + hk) * Wk - kx]));
        }
    }
    kout[idx] = acc;
}

// Copy a wrapped k-image (rows ky=-N/2..N/2-1, row-major) into column-major DFT order.
__global__ void __launch_bounds__(kThreads)
irfft_colfft_f64(const cplx<double>* __restrict__ kin, int N, const cplx<double>* __restrict__ tw,
                 cplx<double>* __restrict__ scratch)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    const int half = N / 2;
    const int W = half + 1;
    const int cpg = cols_per_group(N);
    const int c0 = blockIdx.x * cpg;
    const int ncol = min(cpg, W - c0);
    if (ncol <= 0) return;
    cplx<double>* bufa = lds;
    cplx<double>* bufb = lds + cpg * N;
    for (int e = threadIdx.x; e < ncol * N; e += blockDim.x) {
        const int cl = e / N;
        const int r = e - cl * N;                       // DFT index: ky = r mod N
        const int row = (r < half) ? r + half : r - half;
        bufa[cl * N + r] = kin[row * W + c0 + cl];
    }
    __syncthreads();
    cplx<double>* res = fft_lds<double>(bufa, bufb, N, ncol, tw, 1);
    for (int e = threadIdx.x; e < ncol * N; e += blockDim.x) {
        const int row = e / ncol;                       // output position p = row - N/2
        const int cl = e - row * ncol;
        const int y = (row - half + N) % N;
        scratch[row * W + c0 + cl] = res[cl * N + y];
    }
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
namespace {
thread_local std::string g_err;

template <typename T>
struct TwiddleCache {
    std::mutex m;
    struct Entry { int device; int N; cplx<T>* ptr; };
    std::vector<Entry> entries;
    const cplx<T>* get(int N)
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return nullptr;
        std::lock_guard<std::mutex> lock(m);
        for (auto& e : entries)
            if (e.device == dev && e.N == N) return e.ptr;
        std::vector<cplx<T>> h(N);
        for (int k = 0; k < N; ++k) {
            long double ang = 2.0L * 3.14159265358979323846264338327950288L * k / N;
            h[k].x = T(std::cos(ang));
            h[k].y = T(std::sin(ang));
        }
        cplx<T>* d = nullptr;
        if (hipMalloc(&d, N * sizeof(cplx<T>)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), N * sizeof(cplx<T>), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d);
            return nullptr;
        }
        entries.push_back({dev, N, d});
        return d;
    }
};
TwiddleCache<double>& twd() { static TwiddleCache<double> c; return c; }
TwiddleCache<float>& twf() { static TwiddleCache<float> c; return c; }

// 2^k or 3*2^k with both in-LDS transforms within 64 KiB per workgroup (N <= 2048 in FP64,
// N <= 4096 in FP32).
bool fft_size_ok(int N, size_t elem)
{
    if (N < 4 || N % 2) return false;
    int m = N;
    while (m % 2 == 0) m /= 2;
    if (m != 1 && m != 3) return false;
    return 2 * size_t(cols_per_group(N)) * N * elem <= 65536;
}

int hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return GSMI_E_HIP;
    }
    return GSMI_OK;
}

template <typename T>
int fold_impl(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
              const gsmi_leaf* leaves, Tables<T> tb, void* scratch, void* stream)
{
    if (!fft_size_ok(N, sizeof(cplx<T>))) {
        set_error("unsupported FFT size " + std::to_string(N) +
                  " (need 2^k or 3*2^k, <= 2048 in FP64, <= 4096 in FP32)");
        return GSMI_E_SIZE;
    }
    if (nstamps <= 0) return GSMI_OK;
    const cplx<T>* tw = (sizeof(T) == 8) ? reinterpret_cast<const cplx<T>*>(twd().get(N))
                                         : reinterpret_cast<const cplx<T>*>(twf().get(N));
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    const int cpg = cols_per_group(N);
    hipLaunchKernelGGL(axis_factors<T>, dim3(nstamps), dim3(kThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), stamps, ops, leaves,
                       reinterpret_cast<cplx<T>*>(scratch));
    int rc = hip_check(hipGetLastError(), "axis_factors launch");
    if (rc) return rc;
    dim3 g1((N / 2 + 1 + cpg - 1) / cpg, nstamps);
    size_t lds1 = 2 * size_t(cpg) * N * sizeof(cplx<T>);
    hipLaunchKernelGGL(fold_colfft<T>, g1, dim3(kThreads), lds1,
                       reinterpret_cast<hipStream_t>(stream), stamps, ops, leaves, tb, tw,
                       reinterpret_cast<cplx<T>*>(scratch));
    return hip_check(hipGetLastError(), "fold_colfft launch");
}

template <typename T>
int rows_impl(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, const void* scratch,
              T* out, void* stream)
{
    if (!fft_size_ok(N, sizeof(cplx<T>))) {
        set_error("unsupported FFT size " + std::to_string(N));
        return GSMI_E_SIZE;
    }
    if (nstamps <= 0) return GSMI_OK;
    const cplx<T>* tw = (sizeof(T) == 8) ? reinterpret_cast<const cplx<T>*>(twd().get(N))
                                         : reinterpret_cast<const cplx<T>*>(twf().get(N));
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    const int rpg = rows_per_group(N);
    dim3 g2((max_ny + rpg - 1) / rpg, nstamps);
    size_t lds2 = 2 * size_t(rpg) * (N / 2) * sizeof(cplx<T>);
    hipLaunchKernelGGL(row_c2r<T>, g2, dim3(kThreads), lds2, reinterpret_cast<hipStream_t>(stream),
                       stamps, reinterpret_cast<const cplx<T>*>(scratch), tw, out);
    return hip_check(hipGetLastError(), "row_c2r launch");
}

template <typename T>
int draw_fft_impl(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
                  const gsmi_leaf* leaves, Tables<T> tb, void* scratch, T* out, void* stream,
                  int max_ny)
{
    int rc = fold_impl<T>(stamps, nstamps, N, ops, leaves, tb, scratch, stream);
    if (rc) return rc;
    return rows_impl<T>(stamps, nstamps, N, max_ny, scratch, out, stream);
}
}  // namespace

void set_error(const std::string& msg) { g_err = msg; }

}  // namespace gsmi

using namespace gsmi;

extern "C" {

const char* gsmi_last_error(void) { return g_err.c_str(); }
int gsmi_version(void) { return 1; }

int gsmi_draw_fft_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny,
                      const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                      const int32_t* cliprows, void* scratch, double* out, void* stream)
{
    if (!stamps || !ops || !leaves || !scratch || !out || max_ny <= 0) {
        set_error("gsmi_draw_fft_f64: null buffer or empty stamp");
        return GSMI_E_ARG;
    }
    Tables<double> tb{tables, cliprows};
    return draw_fft_impl<double>(stamps, nstamps, N, ops, leaves, tb, scratch, out, stream, max_ny);
}

int gsmi_draw_fft_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny,
                      const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                      const float* tables32, void* scratch, float* out, void* stream)
{
    if (!stamps || !ops || !leaves || !scratch || !out || max_ny <= 0) {
        set_error("gsmi_draw_fft_f32: null buffer or empty stamp");
        return GSMI_E_ARG;
    }
    Tables<float> tb{tables, tables32};
    return draw_fft_impl<float>(stamps, nstamps, N, ops, leaves, tb, scratch, out, stream, max_ny);
}

int gsmi_draw_kimage_f64(const gsmi_op* ops, int nops, const gsmi_leaf* leaves,
                         const double* tables, const int32_t* cliprows, int xmin, int ymin,
                         int nx, int ny, int Nk, double dk, double* out, void* stream)
{
    if (!ops || !leaves || !out || nx <= 0 || ny <= 0 || nops <= 0) {
        set_error("gsmi_draw_kimage_f64: bad arguments");
        return GSMI_E_ARG;
    }
    Tables<double> tb{tables, cliprows};
    int total = nx * ny;
    hipLaunchKernelGGL(draw_kimage_f64, dim3((total + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), ops, nops, leaves, tb, xmin, ymin,
                       nx, ny, dk, Nk, reinterpret_cast<cplx<double>*>(out));
    return hip_check(hipGetLastError(), "draw_kimage launch");
}

int gsmi_wrap_hermx_f64(const double* kin, int Nk, int N, double* kout, void* stream)
{
    if (!kin || !kout || N < 2 || N % 2 || Nk < N || Nk % 2) {
        set_error("gsmi_wrap_hermx_f64: need even N <= Nk");
        return GSMI_E_ARG;
    }
    int total = (N / 2 + 1) * N;
    hipLaunchKernelGGL(wrap_hermx_f64, dim3((total + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const cplx<double>*>(kin), Nk, N,
                       reinterpret_cast<cplx<double>*>(kout));
    return hip_check(hipGetLastError(), "wrap_hermx launch");
}

int gsmi_irfft_f64(const double* kin, int N, double* out, void* scratch, void* stream)
{
    if (!kin || !out || !scratch) {
        set_error("gsmi_irfft_f64: null buffer");
        return GSMI_E_ARG;
    }
    if (!fft_size_ok(N, sizeof(cplx<double>))) {
        set_error("unsupported FFT size " + std::to_string(N) + " (need 2^k or 3*2^k, <= 2048)");
        return GSMI_E_SIZE;
    }
    const cplx<double>* tw = twd().get(N);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int cpg = cols_per_group(N);
    size_t lds1 = 2 * size_t(cpg) * N * sizeof(cplx<double>);
    hipLaunchKernelGGL(irfft_colfft_f64, dim3((N / 2 + cpg) / cpg), dim3(kThreads), lds1, s,
                       reinterpret_cast<const cplx<double>*>(kin), N, tw,
                       reinterpret_cast<cplx<double>*>(scratch));
    int rc = hip_check(hipGetLastError(), "irfft columns launch");
    if (rc) return rc;
    const int rpg = rows_per_group(N);
    size_t lds2 = 2 * size_t(rpg) * (N / 2) * sizeof(cplx<double>);
    hipLaunchKernelGGL(irfft_rows_f64, dim3((N + rpg - 1) / rpg), dim3(kThreads), lds2, s,
                       reinterpret_cast<const cplx<double>*>(scratch), N, tw, out);
    return hip_check(hipGetLastError(), "irfft rows launch");
}


int gsmi_fold_colfft_f64(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
                         const gsmi_leaf* leaves, const double* tables, const int32_t* cliprows,
                         void* scratch, void* stream)
{
    if (!stamps || !ops || !leaves || !scratch) { set_error("null buffer"); return GSMI_E_ARG; }
    Tables<double> tb{tables, cliprows};
    return fold_impl<double>(stamps, nstamps, N, ops, leaves, tb, scratch, stream);
}

int gsmi_fold_colfft_f32(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
                         const gsmi_leaf* leaves, const double* tables, const float* tables32,
                         void* scratch, void* stream)
{
    if (!stamps || !ops || !leaves || !scratch) { set_error("null buffer"); return GSMI_E_ARG; }
    Tables<float> tb{tables, tables32};
    return fold_impl<float>(stamps, nstamps, N, ops, leaves, tb, scratch, stream);
}

int gsmi_row_c2r_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, const void* scratch,
                     double* out, void* stream)
{
    if (!stamps || !scratch || !out || max_ny <= 0) { set_error("bad argument"); return GSMI_E_ARG; }
    return rows_impl<double>(stamps, nstamps, N, max_ny, scratch, out, stream);
}

int gsmi_row_c2r_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, const void* scratch,
                     float* out, void* stream)
{
    if (!stamps || !scratch || !out || max_ny <= 0) { set_error("bad argument"); return GSMI_E_ARG; }
    return rows_impl<float>(stamps, nstamps, N, max_ny, scratch, out, stream);
}

}
