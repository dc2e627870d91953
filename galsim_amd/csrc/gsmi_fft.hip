// Two-dimensional discrete Fourier transforms of images (galsim.fft and Image.calculate_fft /
// calculate_inverse_fft): complex <-> complex, real -> half-complex and half-complex -> real, with
// the reference's shift_in / shift_out conventions (src/Image.cpp:729-1000, whose FFTW calls these
// replace), on gfx950.
//
// A transform is two passes over lines held in LDS, one workgroup per group of lines:
//   * rows (x, contiguous) and columns (y, stride = row width: a workgroup takes adjacent columns,
//     so each row of its tile is one contiguous run);
//   * the real transforms use the half-length complex transform of z(m) = x(2m) + i x(2m+1) and
//     split its result into the even / odd halves (forward), or build that sequence from the
//     Hermitian half-plane (inverse);
//   * a power-of-two line length from 128 to 4096 runs register-resident (radix-8 passes in
//     registers, one LDS line for the exchanges between them, compile-time indexing);
//   * any other length whose prime factors are 2, 3 and 5 runs as a mixed-radix Stockham
//     transform in LDS (radix 5, 3, 8, 4, 2); any other length L as Bluestein's chirp-z
//     convolution, in LDS, with a 5-smooth transform length S >= 2L - 1;
//   * fftshift of an even axis is a rotation of its index by half its length: an input shift is
//     applied where the line is read, an output shift where it is written (the reference
//     multiplies by the equivalent alternating signs).
// The inverse transform is e^{+2 pi i} (the LDS transforms' own sign); the forward one conjugates
// on the way in and out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsmi.h"
#include "device_program.h"
#include "launch.h"

namespace gsmi {
namespace {

// How one line length is transformed in LDS.
template <typename T>
struct LinePlan {
    int L;                    // transform length
    int S;                    // LDS elements per line: L, or the Bluestein length
    const cplx<T>* tw;        // exp(+2 pi i m / S), m < S
    const cplx<T>* chirp;     // Bluestein only (else nullptr): exp(+i pi m^2 / L), m < L
    const cplx<T>* bh;        // Bluestein only: DFT_S(b) / S with b_m = conj(chirp_|m|), m mod S
};

// Mixed-radix transform of nseg segments of any 5-smooth length n.
template <typename T>
__device__ cplx<T>* fft_lds_any(cplx<T>* a, cplx<T>* b, int n, int nseg, const cplx<T>* tw)
{
    int Ns = 1;
    int rem = n;
    cplx<T>* src = a;
    cplx<T>* dst = b;
    while (rem > 1) {
        int R;
        if (rem % 5 == 0) { stockham_pass<T, 5>(src, dst, n, nseg, Ns, tw, 1); R = 5; }
        else if (rem % 3 == 0) { stockham_pass<T, 3>(src, dst, n, nseg, Ns, tw, 1); R = 3; }
        else if (rem % 8 == 0) { stockham_pass<T, 8>(src, dst, n, nseg, Ns, tw, 1); R = 8; }
        else if (rem % 4 == 0) { stockham_pass<T, 4>(src, dst, n, nseg, Ns, tw, 1); R = 4; }
        else { stockham_pass<T, 2>(src, dst, n, nseg, Ns, tw, 1); R = 2; }
        Ns *= R;
        rem /= R;
        __syncthreads();
        cplx<T>* t = src; src = dst; dst = t;
    }
    return src;
}

// Y(k) = sum_n a(li * S + n) e^{+2 pi i n k / L} for every line li < nl, k < L; the input occupies
// the first L entries of each line of a.  Returns the buffer (a or b) holding Y.  Bluestein:
// 2 n k = n^2 + k^2 - (k - n)^2, so Y(k) = c(k) sum_n (x(n) c(n)) conj(c(k - n)), c(m) =
// e^{+i pi m^2 / L}: a circular convolution of length S >= 2L - 1, done as
// conj(G(conj(G(x c) . G(b) / S))) with G the e^{+2 pi i} transform of length S.
template <typename T>
__device__ cplx<T>* line_dft(cplx<T>* a, cplx<T>* b, int nl, const LinePlan<T>& p)
{
    if (!p.chirp) return fft_lds_any<T>(a, b, p.L, nl, p.tw);
    const int L = p.L, S = p.S;
    const float inv_s = 1.f / float(S);
    for (int e = threadIdx.x; e < nl * S; e += blockDim.x) {
        const int n = e - S * fdiv(e, inv_s);
        a[e] = n < L ? cmul(a[e], p.chirp[n]) : cplx<T>{T(0), T(0)};
    }
    __syncthreads();
    cplx<T>* r = fft_lds_any<T>(a, b, S, nl, p.tw);
    cplx<T>* o = (r == a) ? b : a;
    for (int e = threadIdx.x; e < nl * S; e += blockDim.x) {
        const int k = e - S * fdiv(e, inv_s);
        r[e] = cconj(cmul(r[e], p.bh[k]));
    }
    __syncthreads();
    cplx<T>* y = fft_lds_any<T>(r, o, S, nl, p.tw);
    for (int e = threadIdx.x; e < nl * S; e += blockDim.x) {
        const int k = e - S * fdiv(e, inv_s);
        if (k < L) y[e] = cmul(cconj(y[e]), p.chirp[k]);
    }
    __syncthreads();
    return y;
}

// Where element n of line l lives: (l / D) A + (l % D) B + n es.  A plain set of lines with line
// stride ls is D >= nlines, B = ls; the four-step passes below use the two-level form.
struct LineMap {
    long long D, A, B, es;
    __device__ __forceinline__ long long base(int l, int nlines) const
    {
        return (D >= nlines) ? l * B : (l / D) * A + (l % D) * B;
    }
};

// Twiddles applied to a four-step first pass: output k of line l is multiplied by
// tw[k * n2], n2 = (l / div) % mod, tw = exp(+2 pi i m / L) of the whole length (tw == nullptr:
// none).
template <typename T>
struct PassTwiddle {
    const cplx<T>* tw;
    int div, mod;
};

// Complex transform of `nlines` lines of length L (see LineMap).  Reads line (l + lrot) mod
// nlines at element (n + erot) mod L, writes line l at element k the value Y((k + krot) mod L),
// times scale; conj_io makes it the forward (e^{-2 pi i}) transform (conjugated input, and output
// conjugated after the optional twiddle).  grid: ceil(nlines / nl); dynamic LDS: 2 nl S complex.
template <typename T>
__global__ void __launch_bounds__(kThreads)
c2c_lines(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst, int nlines, int nl,
          LineMap in, int lrot, int erot, LineMap out, int krot, int conj_in, int conj_out,
          T scale, LinePlan<T> p, PassTwiddle<T> ptw)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* bufa = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* bufb = bufa + size_t(nl) * p.S;
    const int L = p.L, S = p.S;
    const bool along_in = (in.es == 1);     // element index fastest in memory
    const bool along_out = (out.es == 1);
    // a tile of lines adjacent in memory reads nl neighbouring values per element: neighbouring
    // tiles go to one XCD so the rest of each cache line is read from its L2
    const int l0 = (along_in ? int(blockIdx.x) : xcd_contiguous(blockIdx.x, gridDim.x)) * nl;
    const int nlc = min(nl, nlines - l0);
    if (nlc <= 0) return;
    const float inv_l = 1.f / float(L);
    const float inv_nl = 1.f / float(nlc);
    for (int e = threadIdx.x; e < nlc * L; e += blockDim.x) {
        int li, n;
        if (along_in) { li = fdiv(e, inv_l); n = e - li * L; }
        else { n = fdiv(e, inv_nl); li = e - n * nlc; }
        int sl = l0 + li + lrot;
        if (sl >= nlines) sl -= nlines;
        int sn = n + erot;
        if (sn >= L) sn -= L;
        cplx<T> v = src[in.base(sl, nlines) + sn * in.es];
        if (conj_in) v.y = -v.y;
        bufa[li * S + n] = v;
    }
    __syncthreads();
    const cplx<T>* y = line_dft<T>(bufa, bufb, nlc, p);
    for (int e = threadIdx.x; e < nlc * L; e += blockDim.x) {
        int li, k;
        if (along_out) { li = fdiv(e, inv_l); k = e - li * L; }
        else { k = fdiv(e, inv_nl); li = e - k * nlc; }
        // output position k holds Y((k + krot) mod L): the rotation by L/2 is its own inverse
        int sk = k + krot;
        if (sk >= L) sk -= L;
        cplx<T> v = y[li * S + sk];
        if (ptw.tw) v = cmul(v, ptw.tw[sk * (((l0 + li) / ptw.div) % ptw.mod)]);
        v.x *= scale;
        v.y *= conj_out ? -scale : scale;
        dst[out.base(l0 + li, nlines) + k * out.es] = v;
    }
}

// Forward real transform along x of an M x N real image into M x (N/2 + 1): line l of the output
// is the transform of input row (l + lrot) mod M read from column erot on (cyclically).  p is the
// plan of length N/2, twn = exp(+2 pi i m / N).  With z(m) = x(2m) + i x(2m+1) and Z = DFT(z):
// X(k) = (Z(k) + conj Z(h-k)) / 2 - i e^{-2 pi i k / N} (Z(k) - conj Z(h-k)) / 2, h = N/2.
template <typename T>
__global__ void __launch_bounds__(kThreads)
r2c_rows(const T* __restrict__ src, cplx<T>* __restrict__ dst, int M, int N, int nl, int lrot,
         int erot, LinePlan<T> p, const cplx<T>* __restrict__ twn)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* bufa = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* bufb = bufa + size_t(nl) * p.S;
    const int h = N / 2, W = h + 1, S = p.S;
    const int l0 = blockIdx.x * nl;
    const int nlc = min(nl, M - l0);
    if (nlc <= 0) return;
    const float inv_h = 1.f / float(h);
    for (int e = threadIdx.x; e < nlc * h; e += blockDim.x) {
        const int li = fdiv(e, inv_h);
        const int m = e - li * h;
        int sl = l0 + li + lrot;
        if (sl >= M) sl -= M;
        const T* row = src + size_t(sl) * N;
        int q = 2 * m + erot;               // x(2m) and x(2m+1) of the rotated row
        if (q >= N) q -= N;
        int q1 = q + 1;
        if (q1 >= N) q1 -= N;
        bufa[li * S + m] = {row[q], -row[q1]};   // conj(z): forward transform
    }
    __syncthreads();
    const cplx<T>* y = line_dft<T>(bufa, bufb, nlc, p);   // conj(Z)
    const float inv_w = 1.f / float(W);
    for (int e = threadIdx.x; e < nlc * W; e += blockDim.x) {
        const int li = fdiv(e, inv_w);
        const int k = e - li * W;
        const int ka = (k == h) ? 0 : k;
        const int kb = (k == 0) ? 0 : h - k;
        const cplx<T> zk = cconj(y[li * S + ka]);
        const cplx<T> zm = cconj(y[li * S + kb]);
        const cplx<T> E = cadd(zk, cconj(zm));
        const cplx<T> O = csub(zk, cconj(zm));
        const cplx<T> t = cmul(cconj(twn[k]), O);
        dst[size_t(l0 + li) * W + k] = {T(0.5) * (E.x + t.y), T(0.5) * (E.y - t.x)};
    }
}

// Inverse real transform along x of an M x (N/2 + 1) half-plane (already transformed along y)
// into an M x N real image: output row l comes from input row (l + lrot) mod M and output
// column j holds x((j + krot) mod N), times scale.  x(2m) = Re v(m), x(2m+1) = Im v(m) with
// v = IDFT_{N/2}(E + i O), E(c) = Z(c) + conj Z(h-c), O(c) = (Z(c) - conj Z(h-c)) e^{2 pi i c / N};
// the imaginary parts of Z(0) and Z(h) are dropped (as numpy.fft.irfft2 does).
template <typename T>
__global__ void __launch_bounds__(kThreads)
c2r_rows(const cplx<T>* __restrict__ src, T* __restrict__ dst, int M, int N, int nl, int lrot,
         int krot, T scale, LinePlan<T> p, const cplx<T>* __restrict__ twn)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* bufa = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* bufb = bufa + size_t(nl) * p.S;
    const int h = N / 2, W = h + 1, S = p.S;
    const int l0 = blockIdx.x * nl;
    const int nlc = min(nl, M - l0);
    if (nlc <= 0) return;
    const float inv_h = 1.f / float(h);
    for (int e = threadIdx.x; e < nlc * h; e += blockDim.x) {
        const int li = fdiv(e, inv_h);
        const int c = e - li * h;
        int sl = l0 + li + lrot;
        if (sl >= M) sl -= M;
        const cplx<T>* z = src + size_t(sl) * W;
        cplx<T> zc = z[c];
        cplx<T> zm = z[h - c];
        if (c == 0) { zc.y = T(0); zm.y = T(0); }
        const cplx<T> E = cadd(zc, cconj(zm));
        const cplx<T> O = cmul(csub(zc, cconj(zm)), twn[c]);
        bufa[li * S + c] = {E.x - O.y, E.y + O.x};
    }
    __syncthreads();
    const cplx<T>* v = line_dft<T>(bufa, bufb, nlc, p);
    const float inv_n = 1.f / float(N);
    for (int e = threadIdx.x; e < nlc * N; e += blockDim.x) {
        const int li = fdiv(e, inv_n);
        const int j = e - li * N;
        int q = j + krot;
        if (q >= N) q -= N;
        const cplx<T> w = v[li * S + (q >> 1)];
        dst[size_t(l0 + li) * N + j] = ((q & 1) ? w.y : w.x) * scale;
    }
}

// ---------------------------------------------------------------------------------------------
// Power-of-two lines of 128..4096 points, register-resident.  The T = L/8 threads of a line hold
// its values eight to a thread (thread t: elements t + s T, s < 8).  A radix pass (8, and a last
// 4 or 2 where L is not a power of 8) transforms a thread's own values in registers; between
// passes the values go through one padded LDS line (written in Stockham order, read back in
// thread order), and the first pass reads and the last pass writes global memory straight from
// registers: two to three LDS round trips per line instead of one per pass plus the staging, one
// LDS line per line instead of two, and every index a compile-time shift or mask.
// ---------------------------------------------------------------------------------------------
// values per thread: 8 for a power-of-two length, 12 for 3 x 2^k (radix 12, then 4 and 2)
template <int L>
constexpr int kRegE = (L % 3 == 0) ? 12 : 8;

// LDS slot of element p: one slot of padding after every 8 values (128 bytes), so the stride-8
// scatter of the first pass falls on distinct banks.
__device__ __forceinline__ int rpad(int p) { return p + (p >> 3); }

// Slots per LDS line: odd, so neighbouring lines of a workgroup start on different banks.
template <int L>
constexpr int kRegLine = L + L / 8 + 1;

// x mod n for 0 <= x < 2n.
__device__ __forceinline__ int wrap_once(int x, int n) { return x >= n ? x - n : x; }

// 16-byte loads and stores of complex values (the kernels below are launched only on 16-byte
// aligned buffers).
__device__ __forceinline__ cplx<double> ld2(const cplx<double>* p)
{
    const double2 x = *reinterpret_cast<const double2*>(p);
    return {x.x, x.y};
}
__device__ __forceinline__ void st2(cplx<double>* p, cplx<double> v)
{
    *reinterpret_cast<double2*>(p) = double2{v.x, v.y};
}

template <int R, typename T>
__device__ __forceinline__ void bfly_reg(cplx<T>* v)
{
    if constexpr (R == 2) bfly2(v);
    if constexpr (R == 4) bfly4(v);
    if constexpr (R == 8) bfly8(v);
    if constexpr (R == 12) bfly12(v);
}

// The radix of the pass after passes whose radices multiply to NS: 8 while at least 8 points
// remain, then the remaining 4 or 2 (E = 8); 12 first, then 4, and a last 2 (E = 12).
template <int L, int NS>
constexpr int reg_radix()
{
    return (kRegE<L> == 12) ? (NS == 1 ? 12 : (L / NS >= 4 ? 4 : 2))
                            : (L / NS >= 8 ? 8 : L / NS);
}

// exp(+2 pi i m / L), m < L <= 4096, from the two 64-entry LDS tables tws[i] = exp(+2 pi i i / L)
// and tws[64 + i] = exp(+2 pi i 64 i / L): one complex product (the passes then wait on no global
// load, so a prefetch of the next lines stays in flight through them).
template <int L>
__device__ __forceinline__ cplx<double> reg_twiddle(const cplx<double>* tws, int m)
{
    if constexpr (L <= 64) return ld2(tws + m);
    else return cmul(ld2(tws + 64 + (m >> 6)), ld2(tws + (m & 63)));
}

// Fill the twiddle tables of reg_twiddle from tw[m] = exp(+2 pi i m / L); ends with a barrier.
template <int L>
__device__ __forceinline__ void reg_twiddle_stage(const cplx<double>* __restrict__ tw,
                                                  cplx<double>* tws)
{
    for (int e = threadIdx.x; e < 128; e += blockDim.x) {
        const int m = e < 64 ? e : (e - 64) * 64;
        st2(tws + e, m < L ? tw[m] : cplx<double>{1., 0.});
    }
    __syncthreads();
}

// Radix-R Stockham pass after passes whose radices multiply to NS, over thread t's values
// (v[s] = element t + s T).  Butterfly j = t + b T (b < 8/R) takes elements j + r L/R, i.e.
// v[b + r 8/R], and its output r goes to (j / NS) NS R + j % NS + r NS: to the LDS line, or, in the
// last pass (NS R = L, so j / NS = 0), to position t + (b + r 8/R) T, back in v.
// tws: the twiddle tables of reg_twiddle.
template <int L, int R, int NS>
__device__ __forceinline__ void reg_pass(cplx<double> (&v)[kRegE<L>], int t,
                                         const cplx<double>* tws, cplx<double>* line)
{
    constexpr int T = L / kRegE<L>, B = kRegE<L> / R;
    cplx<double> out[kRegE<L>];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = t + b * T;
        cplx<double> u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[b + r * B];
        if constexpr (NS > 1) {
            const int k = (NS & (NS - 1)) == 0 ? (j & (NS - 1)) : j % NS;
            cplx<double> w[R];
#pragma unroll
            for (int r = 1; r < R; ++r) w[r] = reg_twiddle<L>(tws, (L / (NS * R)) * r * k);
#pragma unroll
            for (int r = 1; r < R; ++r) u[r] = cmul(u[r], w[r]);
        }
        bfly_reg<R>(u);
        if constexpr (NS * R == L) {
#pragma unroll
            for (int r = 0; r < R; ++r) out[b + r * B] = u[r];
        } else {
            const int jm = (NS & (NS - 1)) == 0 ? (j & (NS - 1)) : j % NS;
            const int base = (j - jm) * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) st2(line + rpad(base + r * NS), u[r]);
        }
    }
    if constexpr (NS * R == L) {
#pragma unroll
        for (int s = 0; s < kRegE<L>; ++s) v[s] = out[s];
    }
}

// The e^{+2 pi i} transform of the line held in v (thread order in, thread order out: v[s] =
// Y(t + s T)).  Every thread of the workgroup calls it (it holds barriers); `line` is this
// thread's line in LDS, free on entry and on return.
template <int L, int NS = 1>
__device__ __forceinline__ void reg_fft(cplx<double> (&v)[kRegE<L>], int t, const cplx<double>* tws,
                                        cplx<double>* line)
{
    constexpr int R = reg_radix<L, NS>();
    reg_pass<L, R, NS>(v, t, tws, line);
    if constexpr (NS * R < L) {
        constexpr int T = L / kRegE<L>;
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kRegE<L>; ++s) v[s] = ld2(line + rpad(t + s * T));
        __syncthreads();
        reg_fft<L, NS * R>(v, t, tws, line);
    }
}

// c2c_lines for plain line maps (line l at l ls, element n at n es) and a register-resident length
// L: nl lines per workgroup of nl L/E threads (at most 512, and 1024 for two columns of 4096; the
// twelve-value forms are bounded to 512 threads, which leaves them the VGPRs their passes need).
// COLS (es != 1): lanes take adjacent lines first, so each access of a wave covers nl adjacent
// values of a row, and neighbouring workgroups go to one XCD; else lanes take adjacent elements
// of a line.
template <int L, bool COLS>
__global__ void __launch_bounds__(kRegE<L> == 12 ? 512 : 1024)
c2c_lines_reg(const cplx<double>* __restrict__ src, cplx<double>* __restrict__ dst, int nlines,
              int nl, long long ls, long long es, int lrot, int erot, int krot, int conj_in,
              int conj_out, double scale, const cplx<double>* __restrict__ tw)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int T = L / kRegE<L>;
    const int li = COLS ? int(threadIdx.x) % nl : int(threadIdx.x) / T;
    const int t = COLS ? int(threadIdx.x) / nl : int(threadIdx.x) % T;
    const int l = (COLS ? xcd_contiguous(blockIdx.x, gridDim.x) : int(blockIdx.x)) * nl + li;
    const bool live = l < nlines;
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* line = lds + li * kRegLine<L>;
    cplx<double>* tws = lds + nl * kRegLine<L>;
    reg_twiddle_stage<L>(tw, tws);
    cplx<double> v[kRegE<L>];
    if (live) {
        int sl = l + lrot;
        if (sl >= nlines) sl -= nlines;
        const cplx<double>* s = src + sl * ls;
#pragma unroll
        for (int q = 0; q < kRegE<L>; ++q) {
            int n = t + q * T + erot;            // erot < L
            if (n >= L) n -= L;
            cplx<double> x = ld2(s + n * es);
            if (conj_in) x.y = -x.y;
            v[q] = x;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kRegE<L>; ++q) v[q] = {0., 0.};
    }
    reg_fft<L>(v, t, tws, line);
    if (!live) return;
    cplx<double>* d = dst + l * ls;
    const double sy = conj_out ? -scale : scale;
#pragma unroll
    for (int q = 0; q < kRegE<L>; ++q) {
        // output position k holds Y((k + krot) mod L)
        int k = t + q * T - krot;               // krot < L
        if (k < 0) k += L;
        st2(d + k * es, {v[q].x * scale, v[q].y * sy});
    }
}

// r2c_rows for N = 2L (L a power of two) and an even column rotation erot: the row's pairs
// x(2m), x(2m+1) are one 16-byte load.
template <int L>
__global__ void __launch_bounds__(kRegE<L> == 12 ? 512 : 1024)
r2c_rows_reg(const double* __restrict__ src, cplx<double>* __restrict__ dst, int M, int nl,
             int lrot, int erot, const cplx<double>* __restrict__ tw,
             const cplx<double>* __restrict__ twn)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int T = L / kRegE<L>, N = 2 * L, h = L, W = L + 1;
    const int li = int(threadIdx.x) / T;
    const int t = int(threadIdx.x) % T;
    const int l = int(blockIdx.x) * nl + li;
    const bool live = l < M;
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* line = lds + li * kRegLine<L>;
    cplx<double>* tws = lds + nl * kRegLine<L>;
    reg_twiddle_stage<L>(tw, tws);
    cplx<double> v[kRegE<L>];
    if (live) {
        int sl = l + lrot;
        if (sl >= M) sl -= M;
        const double* row = src + size_t(sl) * N;
#pragma unroll
        for (int q = 0; q < kRegE<L>; ++q) {
            const double2 x =
                *reinterpret_cast<const double2*>(row + wrap_once(2 * (t + q * T) + erot, N));
            v[q] = {x.x, -x.y};                  // conj(z): forward transform
        }
    } else {
#pragma unroll
        for (int q = 0; q < kRegE<L>; ++q) v[q] = {0., 0.};
    }
    reg_fft<L>(v, t, tws, line);                // conj(Z)
#pragma unroll
    for (int q = 0; q < kRegE<L>; ++q) st2(line + rpad(t + q * T), v[q]);
    __syncthreads();
    if (!live) return;
    cplx<double>* d = dst + size_t(l) * W;
#pragma unroll
    for (int q = 0; q < kRegE<L>; ++q) {
        const int k = t + q * T;
        const cplx<double> zk = cconj(v[q]);
        const cplx<double> zm = cconj(ld2(line + rpad(k == 0 ? 0 : h - k)));
        const cplx<double> E = cadd(zk, cconj(zm));
        const cplx<double> O = csub(zk, cconj(zm));
        const cplx<double> u = cmul(cconj(twn[k]), O);
        st2(d + k, {0.5 * (E.x + u.y), 0.5 * (E.y - u.x)});
    }
    if (t == 0) {
        // k = h: Z(h) = Z(0)
        const cplx<double> z0 = cconj(v[0]);
        const cplx<double> E = cadd(z0, cconj(z0));
        const cplx<double> O = csub(z0, cconj(z0));
        const cplx<double> u = cmul(cconj(twn[h]), O);
        st2(d + h, {0.5 * (E.x + u.y), 0.5 * (E.y - u.x)});
    }
}

// c2r_rows for N = 2L (L a power of two) and an even column rotation krot: pixels x(2m), x(2m+1)
// are one 16-byte store.
template <int L>
__global__ void __launch_bounds__(kRegE<L> == 12 ? 512 : 1024)
c2r_rows_reg(const cplx<double>* __restrict__ src, double* __restrict__ dst, int M, int nl,
             int lrot, int krot, double scale, const cplx<double>* __restrict__ tw,
             const cplx<double>* __restrict__ twn)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int T = L / kRegE<L>, N = 2 * L, h = L, W = L + 1;
    const int li = int(threadIdx.x) / T;
    const int t = int(threadIdx.x) % T;
    const int l = int(blockIdx.x) * nl + li;
    const bool live = l < M;
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* line = lds + li * kRegLine<L>;
    cplx<double>* tws = lds + nl * kRegLine<L>;
    reg_twiddle_stage<L>(tw, tws);
    cplx<double> v[kRegE<L>];
    if (live) {
        int sl = l + lrot;
        if (sl >= M) sl -= M;
        const cplx<double>* z = src + size_t(sl) * W;
#pragma unroll
        for (int q = 0; q < kRegE<L>; ++q) {
            const int c = t + q * T;
            cplx<double> zc = ld2(z + c);
            cplx<double> zm = ld2(z + (h - c));
            if (c == 0) { zc.y = 0.; zm.y = 0.; }
            const cplx<double> E = cadd(zc, cconj(zm));
            const cplx<double> O = cmul(csub(zc, cconj(zm)), twn[c]);
            v[q] = {E.x - O.y, E.y + O.x};
        }
    } else {
#pragma unroll
        for (int q = 0; q < kRegE<L>; ++q) v[q] = {0., 0.};
    }
    reg_fft<L>(v, t, tws, line);
    if (!live) return;
    double* o = dst + size_t(l) * N;
#pragma unroll
    for (int q = 0; q < kRegE<L>; ++q) {
        // x(2m) and x(2m+1) of m = t + q T land at columns 2m - krot and 2m + 1 - krot (mod N)
        const int j = wrap_once(2 * (t + q * T) - krot + N, N);
        *reinterpret_cast<double2*>(o + j) = double2{v[q].x * scale, v[q].y * scale};
    }
}

// Bluestein's algorithm (L. I. Bluestein, 1970) for a line of any length L with 2L - 1 <= S, S a
// power of two, register-resident: v holds x(n) in thread order (zero for n >= L); on return
// v[q] = Y(k) = sum_n x(n) e^{+2 pi i n k / L} for k = t + q S/8 < L (zero beyond).  With
// 2 n k = n^2 + k^2 - (k - n)^2 and c(m) = e^{+i pi m^2 / L}: Y = c . (conj(G(conj(G(x c) . G(b) /
// S))), G the length-S e^{+2 pi i} transform, b(m mod S) = conj c(|m|); chirp = c(0..L-1),
// bh = G(b) / S.
template <int S>
__device__ __forceinline__ void reg_bluestein(cplx<double> (&v)[kRegE<S>], int t, int L,
                                              const cplx<double>* __restrict__ chirp,
                                              const cplx<double>* __restrict__ bh,
                                              const cplx<double>* tws, cplx<double>* line)
{
    constexpr int T = S / kRegE<S>;
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int n = t + q * T;
        v[q] = n < L ? cmul(v[q], chirp[n]) : cplx<double>{0., 0.};
    }
    reg_fft<S>(v, t, tws, line);
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) v[q] = cconj(cmul(v[q], bh[t + q * T]));
    reg_fft<S>(v, t, tws, line);
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int k = t + q * T;
        v[q] = k < L ? cmul(cconj(v[q]), chirp[k]) : cplx<double>{0., 0.};
    }
}

// c2c_lines for plain line maps and a line length L that takes Bluestein's algorithm (2L - 1 <=
// S): as c2c_lines_reg (nl lines per workgroup of nl S/8 threads; COLS: lanes take adjacent lines
// first), with any rotations.
template <int S, bool COLS>
__global__ void __launch_bounds__(1024)
c2c_lines_blue(const cplx<double>* __restrict__ src, cplx<double>* __restrict__ dst, int nlines,
               int nl, int L, long long ls, long long es, int lrot, int erot, int krot,
               int conj_in, int conj_out, double scale, const cplx<double>* __restrict__ tw,
               const cplx<double>* __restrict__ chirp, const cplx<double>* __restrict__ bh)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int T = S / kRegE<S>;
    const int li = COLS ? int(threadIdx.x) % nl : int(threadIdx.x) / T;
    const int t = COLS ? int(threadIdx.x) / nl : int(threadIdx.x) % T;
    const int l = (COLS ? xcd_contiguous(blockIdx.x, gridDim.x) : int(blockIdx.x)) * nl + li;
    const bool live = l < nlines;
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* line = lds + li * kRegLine<S>;
    cplx<double>* tws = lds + nl * kRegLine<S>;
    reg_twiddle_stage<S>(tw, tws);
    int sl = l + lrot;
    if (sl >= nlines) sl -= nlines;
    const cplx<double>* sp = src + (live ? sl : 0) * ls;
    cplx<double> v[kRegE<S>];
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int n = t + q * T;
        if (live && n < L) {
            int sn = n + erot;
            if (sn >= L) sn -= L;
            cplx<double> x = sp[sn * es];
            if (conj_in) x.y = -x.y;
            v[q] = x;
        } else {
            v[q] = {0., 0.};
        }
    }
    reg_bluestein<S>(v, t, L, chirp, bh, tws, line);
    if (!live) return;
    cplx<double>* d = dst + l * ls;
    const double sy = conj_out ? -scale : scale;
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int kk = t + q * T;            // holds Y(kk), written at position (kk - krot) mod L
        if (kk >= L) continue;
        int k = kk - krot;
        if (k < 0) k += L;
        d[k * es] = {v[q].x * scale, v[q].y * sy};
    }
}

// r2c_rows for a half-length L = N/2 that takes Bluestein's algorithm (2L - 1 <= S), any column
// rotation erot: nl rows per workgroup of nl S/8 threads.
template <int S>
__global__ void __launch_bounds__(1024)
r2c_rows_blue(const double* __restrict__ src, cplx<double>* __restrict__ dst, int M, int N, int nl,
              int lrot, int erot, const cplx<double>* __restrict__ tw,
              const cplx<double>* __restrict__ twn, const cplx<double>* __restrict__ chirp,
              const cplx<double>* __restrict__ bh)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int T = S / kRegE<S>;
    const int h = N / 2, L = h, W = h + 1;
    const int li = int(threadIdx.x) / T;
    const int t = int(threadIdx.x) % T;
    const int l = int(blockIdx.x) * nl + li;
    const bool live = l < M;
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* line = lds + li * kRegLine<S>;
    cplx<double>* tws = lds + nl * kRegLine<S>;
    reg_twiddle_stage<S>(tw, tws);
    cplx<double> v[kRegE<S>];
    int sl = l + lrot;
    if (sl >= M) sl -= M;
    const double* row = src + size_t(live ? sl : 0) * N;
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int m = t + q * T;
        if (live && m < L) {
            int q0 = 2 * m + erot;               // x(2m) and x(2m+1) of the rotated row
            if (q0 >= N) q0 -= N;
            int q1 = q0 + 1;
            if (q1 >= N) q1 -= N;
            v[q] = {row[q0], -row[q1]};          // conj(z): forward transform
        } else {
            v[q] = {0., 0.};
        }
    }
    reg_bluestein<S>(v, t, L, chirp, bh, tws, line);   // conj(Z)
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) line[rpad(t + q * T)] = v[q];
    __syncthreads();
    if (!live) return;
    cplx<double>* d = dst + size_t(l) * W;
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int k = t + q * T;
        if (k >= L) continue;
        const cplx<double> zk = cconj(v[q]);
        const cplx<double> zm = cconj(line[rpad(k == 0 ? 0 : h - k)]);
        const cplx<double> E = cadd(zk, cconj(zm));
        const cplx<double> O = csub(zk, cconj(zm));
        const cplx<double> u = cmul(cconj(twn[k]), O);
        d[k] = {0.5 * (E.x + u.y), 0.5 * (E.y - u.x)};
    }
    if (t == 0) {
        const cplx<double> z0 = cconj(v[0]);             // k = h: Z(h) = Z(0)
        const cplx<double> E = cadd(z0, cconj(z0));
        const cplx<double> O = csub(z0, cconj(z0));
        const cplx<double> u = cmul(cconj(twn[h]), O);
        d[h] = {0.5 * (E.x + u.y), 0.5 * (E.y - u.x)};
    }
}

// c2r_rows for a half-length L = N/2 that takes Bluestein's algorithm (2L - 1 <= S), any column
// rotation krot.
template <int S>
__global__ void __launch_bounds__(1024)
c2r_rows_blue(const cplx<double>* __restrict__ src, double* __restrict__ dst, int M, int N, int nl,
              int lrot, int krot, double scale, const cplx<double>* __restrict__ tw,
              const cplx<double>* __restrict__ twn, const cplx<double>* __restrict__ chirp,
              const cplx<double>* __restrict__ bh)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int T = S / kRegE<S>;
    const int h = N / 2, L = h, W = h + 1;
    const int li = int(threadIdx.x) / T;
    const int t = int(threadIdx.x) % T;
    const int l = int(blockIdx.x) * nl + li;
    const bool live = l < M;
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* line = lds + li * kRegLine<S>;
    cplx<double>* tws = lds + nl * kRegLine<S>;
    reg_twiddle_stage<S>(tw, tws);
    cplx<double> v[kRegE<S>];
    int sl = l + lrot;
    if (sl >= M) sl -= M;
    const cplx<double>* z = src + size_t(live ? sl : 0) * W;
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int c = t + q * T;
        if (live && c < L) {
            cplx<double> zc = z[c];
            cplx<double> zm = z[h - c];
            if (c == 0) { zc.y = 0.; zm.y = 0.; }
            const cplx<double> E = cadd(zc, cconj(zm));
            const cplx<double> O = cmul(csub(zc, cconj(zm)), twn[c]);
            v[q] = {E.x - O.y, E.y + O.x};
        } else {
            v[q] = {0., 0.};
        }
    }
    reg_bluestein<S>(v, t, L, chirp, bh, tws, line);
    if (!live) return;
    double* o = dst + size_t(l) * N;
#pragma unroll
    for (int q = 0; q < kRegE<S>; ++q) {
        const int m = t + q * T;
        if (m >= L) continue;
        // x(2m) and x(2m+1) land at columns 2m - krot and 2m + 1 - krot (mod N)
        int j0 = 2 * m - krot;
        if (j0 < 0) j0 += N;
        int j1 = j0 + 1;
        if (j1 >= N) j1 -= N;
        o[j0] = v[q].x * scale;
        o[j1] = v[q].y * scale;
    }
}

// The full M x N transform of a real image from its M x (N/2 + 1) half (galsim/fft.py fft2's
// Hermitian fill): column j >= N/2 (or j < N/2 with shift_out, whose columns are rotated by N/2)
// is conj of the half-plane at row (M - i) mod M, column N - j (N/2 - j with shift_out).
template <typename T>
__global__ void __launch_bounds__(kThreads)
hermitian_fill(const cplx<T>* __restrict__ rk, int M, int N, int shift_out,
               cplx<T>* __restrict__ out)
{
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= size_t(M) * N) return;
    const int i = int(idx / N);
    const int j = int(idx - size_t(i) * N);
    const int h = N / 2, W = h + 1;
    const int im = (i == 0) ? 0 : M - i;
    cplx<T> v;
    if (!shift_out) {
        v = (j < h) ? rk[size_t(i) * W + j] : cconj(rk[size_t(im) * W + (N - j)]);
    } else {
        v = (j >= h) ? rk[size_t(i) * W + (j - h)] : cconj(rk[size_t(im) * W + (h - j)]);
    }
    out[idx] = v;
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
bool smooth5(int n)
{
    if (n < 1) return false;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

int next_smooth5(int n)
{
    while (!smooth5(n)) ++n;
    return n;
}

// Bluestein tables per (device, L): chirp c(m) = exp(+i pi (m^2 mod 2L) / L) and
// bh = DFT_S(b) / S, b(m mod S) = conj c(|m|) for |m| < L, computed in long double.
template <typename T>
struct ChirpCache {
    std::mutex m;
    struct Entry { int device; int L; int S; cplx<T>* chirp; cplx<T>* bh; };
    std::vector<Entry> entries;
    bool get(int L, int S, const cplx<T>** chirp, const cplx<T>** bh)
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        std::lock_guard<std::mutex> lock(m);
        for (auto& e : entries)
            if (e.device == dev && e.L == L && e.S == S) {
                *chirp = e.chirp;
                *bh = e.bh;
                return true;
            }
        const long double pi = 3.14159265358979323846264338327950288L;
        std::vector<long double> cr(L), ci(L);
        for (int k = 0; k < L; ++k) {
            long long q = (static_cast<long long>(k) * k) % (2LL * L);
            cr[k] = std::cos(pi * q / L);
            ci[k] = std::sin(pi * q / L);
        }
        std::vector<long double> br(S, 0.L), bi(S, 0.L);
        for (int k = 0; k < L; ++k) {
            br[k] = cr[k];
            bi[k] = -ci[k];
            if (k > 0) {
                br[S - k] = cr[k];
                bi[S - k] = -ci[k];
            }
        }
        std::vector<long double> rr(S), ri(S);
        for (int k = 0; k < S; ++k) {
            rr[k] = std::cos(2.L * pi * k / S);
            ri[k] = std::sin(2.L * pi * k / S);
        }
        std::vector<cplx<T>> hc(L), hb(S);
        for (int k = 0; k < L; ++k) hc[k] = {T(cr[k]), T(ci[k])};
        for (int k = 0; k < S; ++k) {
            long double sr = 0.L, si = 0.L;
            long long idx = 0;
            for (int m = 0; m < S; ++m, idx += k) {
                if (idx >= S) idx %= S;
                sr += br[m] * rr[idx] - bi[m] * ri[idx];
                si += br[m] * ri[idx] + bi[m] * rr[idx];
            }
            hb[k] = {T(sr / S), T(si / S)};
        }
        cplx<T>* dc = nullptr;
        cplx<T>* db = nullptr;
        if (hipMalloc(&dc, L * sizeof(cplx<T>)) != hipSuccess) return false;
        if (hipMalloc(&db, S * sizeof(cplx<T>)) != hipSuccess) {
            (void)hipFree(dc);
            return false;
        }
        if (hipMemcpy(dc, hc.data(), L * sizeof(cplx<T>), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(db, hb.data(), S * sizeof(cplx<T>), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(dc);
            (void)hipFree(db);
            return false;
        }
        entries.push_back({dev, L, S, dc, db});
        *chirp = dc;
        *bh = db;
        return true;
    }
};
ChirpCache<double>& chirps() { static ChirpCache<double> c; return c; }

// Plan for length L, or GSMI_E_SIZE / GSMI_E_HIP with the error recorded.
int make_plan(int L, size_t lds, LinePlan<double>* p)
{
    const size_t elem = sizeof(cplx<double>);
    p->L = L;
    p->chirp = nullptr;
    p->bh = nullptr;
    if (smooth5(L) && 2 * size_t(L) * elem <= lds) {
        p->S = L;
    } else {
        p->S = next_smooth5(2 * L - 1);
        if (2 * size_t(p->S) * elem > lds) {
            set_error("FFT length " + std::to_string(L) + " exceeds the in-LDS limit (" +
                      std::to_string(lds / (2 * elem)) + " complex per line; lengths with prime "
                      "factors other than 2, 3, 5 need 2L-1 rounded up to such a length)");
            return GSMI_E_SIZE;
        }
        if (!chirps().get(L, p->S, &p->chirp, &p->bh)) {
            set_error("Bluestein table allocation failed");
            return GSMI_E_HIP;
        }
    }
    p->tw = twiddles<double>(p->S);
    if (!p->tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    return GSMI_OK;
}

// Lines per workgroup: about 2048 elements of work, at least one line, within the LDS limit.
int lines_per_group(int S, int nlines, size_t lds)
{
    int nl = std::max(1, 2048 / S);
    nl = std::min(nl, nlines);
    while (nl > 1 && 2 * size_t(nl) * S * sizeof(cplx<double>) > lds) --nl;
    return nl;
}

// The register-resident kernels: power-of-two lengths 128..4096 and 3 x 2^k lengths 192..3072 on
// 16-byte aligned buffers,
// 4096 / L lines (512 threads) per workgroup, two LDS lines (about 74 KiB) and two workgroups
// per CU.
bool reg_len(int L)
{
    const int p = (L % 3 == 0) ? L / 3 : L;
    return L >= 128 && L <= 4096 && (p & (p - 1)) == 0;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Lines per workgroup of the register kernels along rows: 2048/L (at most 256 threads); 4096/L
// measured 1-15% slower (1174^2, 1536^2, 4096^2 rfft2 + irfft2).
int reg_lines(int L, int nlines) { return std::max(1, std::min(nlines, 2048 / L)); }

template <int L>
int launch_c2c_reg_t(const cplx<double>* src, cplx<double>* dst, int nlines, long long ls,
                     long long es, int lrot, int erot, int krot, int conj_io, double scale,
                     hipStream_t s)
{
    const cplx<double>* tw = twiddles<double>(L);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    // columns: 4096/L per workgroup (two of 4096), so a wave's row accesses are 32 bytes or more;
    // one column fewer or twice as many per workgroup both measured slower
    const int nl = (es != 1) ? std::max(1, std::min(nlines, L == 4096 ? 2 : 4096 / L))
                             : reg_lines(L, nlines);
    const size_t bytes = (size_t(nl) * kRegLine<L> + 128) * sizeof(cplx<double>);
    const dim3 grid((nlines + nl - 1) / nl), block(nl * (L / kRegE<L>));
    int rc;
    if (es == 1) {
        if ((rc = allow_lds(reinterpret_cast<const void*>(c2c_lines_reg<L, false>), bytes)))
            return rc;
        hipLaunchKernelGGL((c2c_lines_reg<L, false>), grid, block, bytes, s, src, dst, nlines, nl,
                           ls, es, lrot, erot, krot, conj_io, conj_io, scale, tw);
    } else {
        if ((rc = allow_lds(reinterpret_cast<const void*>(c2c_lines_reg<L, true>), bytes)))
            return rc;
        hipLaunchKernelGGL((c2c_lines_reg<L, true>), grid, block, bytes, s, src, dst, nlines, nl,
                           ls, es, lrot, erot, krot, conj_io, conj_io, scale, tw);
    }
    return hip_check(hipGetLastError(), "c2c_lines_reg launch");
}

template <int L>
int launch_r2c_reg_t(const double* src, cplx<double>* dst, int M, int lrot, int erot,
                     hipStream_t s)
{
    const cplx<double>* tw = twiddles<double>(L);
    const cplx<double>* twn = twiddles<double>(2 * L);
    if (!tw || !twn) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    const int nl = reg_lines(L, M);
    const size_t bytes = (size_t(nl) * kRegLine<L> + 128) * sizeof(cplx<double>);
    int rc = allow_lds(reinterpret_cast<const void*>(r2c_rows_reg<L>), bytes);
    if (rc) return rc;
    hipLaunchKernelGGL(r2c_rows_reg<L>, dim3((M + nl - 1) / nl), dim3(nl * (L / kRegE<L>)), bytes, s,
                       src, dst, M, nl, lrot, erot, tw, twn);
    return hip_check(hipGetLastError(), "r2c_rows_reg launch");
}

template <int L>
int launch_c2r_reg_t(const cplx<double>* src, double* dst, int M, int lrot, int krot,
                     double scale, hipStream_t s)
{
    const cplx<double>* tw = twiddles<double>(L);
    const cplx<double>* twn = twiddles<double>(2 * L);
    if (!tw || !twn) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    const int nl = reg_lines(L, M);
    const size_t bytes = (size_t(nl) * kRegLine<L> + 128) * sizeof(cplx<double>);
    int rc = allow_lds(reinterpret_cast<const void*>(c2r_rows_reg<L>), bytes);
    if (rc) return rc;
    hipLaunchKernelGGL(c2r_rows_reg<L>, dim3((M + nl - 1) / nl), dim3(nl * (L / kRegE<L>)), bytes, s,
                       src, dst, M, nl, lrot, krot, scale, tw, twn);
    return hip_check(hipGetLastError(), "c2r_rows_reg launch");
}

// The register-resident Bluestein length for a line length L that is not 5-smooth: the smallest
// power of two in 128..4096 that is >= 2L - 1, or 0 if there is none (a 3 x 2^k length, though
// shorter, ran slower: its twelve-value passes hold more registers).
int blue_len(int L)
{
    int S = 128;
    while (S < 2 * L - 1) S *= 2;
    return S <= 4096 ? S : 0;
}

template <int S>
int launch_blue_rows_t(bool r2c, const void* src, void* dst, int M, int N, int lrot, int rot,
                       double scale, hipStream_t s)
{
    const int L = N / 2;
    const cplx<double>* tw = twiddles<double>(S);
    const cplx<double>* twn = twiddles<double>(N);
    const cplx<double>* chirp = nullptr;
    const cplx<double>* bh = nullptr;
    if (!tw || !twn) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    if (!chirps().get(L, S, &chirp, &bh)) {
        set_error("Bluestein table allocation failed");
        return GSMI_E_HIP;
    }
    const int nl = reg_lines(S, M);
    const size_t bytes = (size_t(nl) * kRegLine<S> + 128) * sizeof(cplx<double>);
    const dim3 grid((M + nl - 1) / nl), block(nl * (S / kRegE<S>));
    int rc;
    if (r2c) {
        if ((rc = allow_lds(reinterpret_cast<const void*>(r2c_rows_blue<S>), bytes))) return rc;
        hipLaunchKernelGGL(r2c_rows_blue<S>, grid, block, bytes, s,
                           reinterpret_cast<const double*>(src),
                           reinterpret_cast<cplx<double>*>(dst), M, N, nl, lrot, rot, tw, twn,
                           chirp, bh);
    } else {
        if ((rc = allow_lds(reinterpret_cast<const void*>(c2r_rows_blue<S>), bytes))) return rc;
        hipLaunchKernelGGL(c2r_rows_blue<S>, grid, block, bytes, s,
                           reinterpret_cast<const cplx<double>*>(src),
                           reinterpret_cast<double*>(dst), M, N, nl, lrot, rot, scale, tw, twn,
                           chirp, bh);
    }
    return hip_check(hipGetLastError(), "Bluestein row launch");
}

template <int S>
int launch_c2c_blue_t(const cplx<double>* src, cplx<double>* dst, int nlines, int L, long long ls,
                      long long es, int lrot, int erot, int krot, int conj_io, double scale,
                      hipStream_t s)
{
    const cplx<double>* tw = twiddles<double>(S);
    const cplx<double>* chirp = nullptr;
    const cplx<double>* bh = nullptr;
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    if (!chirps().get(L, S, &chirp, &bh)) {
        set_error("Bluestein table allocation failed");
        return GSMI_E_HIP;
    }
    const int nl = reg_lines(S, nlines);
    const size_t bytes = (size_t(nl) * kRegLine<S> + 128) * sizeof(cplx<double>);
    const dim3 grid((nlines + nl - 1) / nl), block(nl * (S / kRegE<S>));
    int rc;
    if (es == 1) {
        if ((rc = allow_lds(reinterpret_cast<const void*>(c2c_lines_blue<S, false>), bytes)))
            return rc;
        hipLaunchKernelGGL((c2c_lines_blue<S, false>), grid, block, bytes, s, src, dst, nlines,
                           nl, L, ls, es, lrot, erot, krot, conj_io, conj_io, scale, tw, chirp,
                           bh);
    } else {
        if ((rc = allow_lds(reinterpret_cast<const void*>(c2c_lines_blue<S, true>), bytes)))
            return rc;
        hipLaunchKernelGGL((c2c_lines_blue<S, true>), grid, block, bytes, s, src, dst, nlines,
                           nl, L, ls, es, lrot, erot, krot, conj_io, conj_io, scale, tw, chirp,
                           bh);
    }
    return hip_check(hipGetLastError(), "Bluestein line launch");
}

int launch_c2c_blue(const cplx<double>* src, cplx<double>* dst, int nlines, int L, long long ls,
                    long long es, int lrot, int erot, int krot, int conj_io, double scale,
                    hipStream_t s)
{
#define GSMI_C2C_BLUE(SS) launch_c2c_blue_t<SS>(src, dst, nlines, L, ls, es, lrot, erot, krot, \
                                                conj_io, scale, s)
    switch (blue_len(L)) {
      case 128: return GSMI_C2C_BLUE(128);
      case 256: return GSMI_C2C_BLUE(256);
      case 512: return GSMI_C2C_BLUE(512);
      case 1024: return GSMI_C2C_BLUE(1024);
      case 2048: return GSMI_C2C_BLUE(2048);
      case 4096: return GSMI_C2C_BLUE(4096);
      default: set_error("no register-resident Bluestein length"); return GSMI_E_SIZE;
    }
#undef GSMI_C2C_BLUE
}

int launch_blue_rows(bool r2c, const void* src, void* dst, int M, int N, int lrot, int rot,
                     double scale, hipStream_t s)
{
    switch (blue_len(N / 2)) {
      case 128: return launch_blue_rows_t<128>(r2c, src, dst, M, N, lrot, rot, scale, s);
      case 256: return launch_blue_rows_t<256>(r2c, src, dst, M, N, lrot, rot, scale, s);
      case 512: return launch_blue_rows_t<512>(r2c, src, dst, M, N, lrot, rot, scale, s);
      case 1024: return launch_blue_rows_t<1024>(r2c, src, dst, M, N, lrot, rot, scale, s);
      case 2048: return launch_blue_rows_t<2048>(r2c, src, dst, M, N, lrot, rot, scale, s);
      case 4096: return launch_blue_rows_t<4096>(r2c, src, dst, M, N, lrot, rot, scale, s);
      default: set_error("no register-resident Bluestein length"); return GSMI_E_SIZE;
    }
}

#define GSMI_REG_SWITCH(L, CALL)                                                                \
    switch (L) {                                                                                \
      case 128: return CALL(128);                                                               \
      case 256: return CALL(256);                                                               \
      case 512: return CALL(512);                                                               \
      case 1024: return CALL(1024);                                                             \
      case 2048: return CALL(2048);                                                             \
      case 4096: return CALL(4096);                                                             \
      case 192: return CALL(192);                                                               \
      case 384: return CALL(384);                                                               \
      case 768: return CALL(768);                                                               \
      case 1536: return CALL(1536);                                                             \
      case 3072: return CALL(3072);                                                             \
      default: set_error("no register-resident transform for this length"); return GSMI_E_SIZE; \
    }

int launch_c2c_reg(int L, const cplx<double>* src, cplx<double>* dst, int nlines, long long ls,
                   long long es, int lrot, int erot, int krot, int conj_io, double scale,
                   hipStream_t s)
{
#define GSMI_C2C(NN) launch_c2c_reg_t<NN>(src, dst, nlines, ls, es, lrot, erot, krot, conj_io, \
                                          scale, s)
    GSMI_REG_SWITCH(L, GSMI_C2C)
#undef GSMI_C2C
}

int launch_r2c_reg(int L, const double* src, cplx<double>* dst, int M, int lrot, int erot,
                   hipStream_t s)
{
#define GSMI_R2C(NN) launch_r2c_reg_t<NN>(src, dst, M, lrot, erot, s)
    GSMI_REG_SWITCH(L, GSMI_R2C)
#undef GSMI_R2C
}

int launch_c2r_reg(int L, const cplx<double>* src, double* dst, int M, int lrot, int krot,
                   double scale, hipStream_t s)
{
#define GSMI_C2R(NN) launch_c2r_reg_t<NN>(src, dst, M, lrot, krot, scale, s)
    GSMI_REG_SWITCH(L, GSMI_C2R)
#undef GSMI_C2R
}
#undef GSMI_REG_SWITCH

int launch_lines(const cplx<double>* src, cplx<double>* dst, int nlines, LineMap in, int lrot,
                 int erot, LineMap out, int krot, int conj_in, int conj_out, double scale,
                 const LinePlan<double>& p, PassTwiddle<double> ptw, size_t lds, hipStream_t s)
{
    const int nl = lines_per_group(p.S, nlines, lds);
    const size_t bytes = 2 * size_t(nl) * p.S * sizeof(cplx<double>);
    int rc = allow_lds(reinterpret_cast<const void*>(c2c_lines<double>), bytes);
    if (rc) return rc;
    hipLaunchKernelGGL(c2c_lines<double>, dim3((nlines + nl - 1) / nl), dim3(kThreads), bytes, s,
                       src, dst, nlines, nl, in, lrot, erot, out, krot, conj_in, conj_out, scale,
                       p, ptw);
    return hip_check(hipGetLastError(), "c2c_lines launch");
}

// One pass over plain lines (line stride ls, element stride es) that fit the LDS.
int launch_c2c(const cplx<double>* src, cplx<double>* dst, int nlines, long long ls, long long es,
               int lrot, int erot, int krot, int conj_io, double scale, const LinePlan<double>& p,
               size_t lds, hipStream_t s)
{
    if (reg_len(p.L) && !p.chirp && aligned16(src) && aligned16(dst))
        return launch_c2c_reg(p.L, src, dst, nlines, ls, es, lrot, erot, krot, conj_io, scale, s);
    if (p.chirp && blue_len(p.L))
        return launch_c2c_blue(src, dst, nlines, p.L, ls, es, lrot, erot, krot, conj_io, scale,
                               s);
    const LineMap m{nlines, 0, ls, es};
    return launch_lines(src, dst, nlines, m, lrot, erot, m, krot, conj_io, conj_io, scale, p,
                        PassTwiddle<double>{nullptr, 1, 1}, lds, s);
}

// Columns too long for one workgroup's LDS: the four-step DFT of length L = L1 L2 (both factors
// fitting): with n = n1 L2 + n2 and k = k1 + L1 k2, first L2 W transforms of length L1 over n1
// (the data read as the n1 x (n2, c) matrix, output times exp(+2 pi i k1 n2 / L), written back in
// place), then L1 W transforms of length L2 over n2 written to k = k1 + L1 k2 in dst (which must
// differ from tmp).  Columns are W adjacent values per row (row stride W); the input rotation
// erot = L/2 (shift) rotates n1 by L1/2 and the output rotation L/2 rotates k2 by L2/2, so L1 and
// L2 are even.  src -> tmp -> dst; tmp may equal src.
struct LongPlan {
    int L1 = 0, L2 = 0;
    LinePlan<double> p1, p2;
    const cplx<double>* tw = nullptr;
};

int make_long_plan(int L, size_t lds, LongPlan* lp)
{
    // the most balanced even factorisation with both factors transformable in LDS
    int best = 0;
    for (int a = 2; a * a <= L; a += 2) {
        if (L % a) continue;
        const int b = L / a;
        if (b % 2) continue;
        best = a;
    }
    if (!best) {
        set_error("FFT length " + std::to_string(L) + " has no even factorisation for the "
                  "four-step transform");
        return GSMI_E_SIZE;
    }
    lp->L1 = best;
    lp->L2 = L / best;
    int rc;
    if ((rc = make_plan(lp->L1, lds, &lp->p1)) || (rc = make_plan(lp->L2, lds, &lp->p2)))
        return rc;
    lp->tw = twiddles<double>(L);
    if (!lp->tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    return GSMI_OK;
}

int launch_columns_long(const cplx<double>* src, cplx<double>* tmp, cplx<double>* dst, int W,
                        const LongPlan& lp, int erot, int krot, int conj_io, double scale,
                        size_t lds, hipStream_t s)
{
    const long long L1 = lp.L1, L2 = lp.L2;
    // pass 1: line l = n2 W + c, element n1 at n1 (L2 W) + l
    const int n1lines = int(L2 * W);
    const LineMap m1{n1lines, 0, 1, L2 * W};
    int rc = launch_lines(src, tmp, n1lines, m1, 0, erot ? int(L1 / 2) : 0, m1, 0, conj_io, 0,
                          1., lp.p1, PassTwiddle<double>{lp.tw, W, int(L2)}, lds, s);
    if (rc) return rc;
    // pass 2: line l = k1 W + c, element n2 at k1 (L2 W) + c + n2 W; output k2 at
    // (k1 + L1 k2) W + c = l + k2 (L1 W)
    const int n2lines = int(L1 * W);
    const LineMap in2{W, L2 * W, 1, W};
    const LineMap out2{n2lines, 0, 1, L1 * W};
    return launch_lines(tmp, dst, n2lines, in2, 0, 0, out2, krot ? int(L2 / 2) : 0, 0, conj_io,
                        scale, lp.p2, PassTwiddle<double>{nullptr, 1, 1}, lds, s);
}

// A column length transformable in one pass, or the four-step plan for it.
int column_plan(int L, size_t lds, LinePlan<double>* p, LongPlan* lp, bool* longp)
{
    const size_t elem = sizeof(cplx<double>);
    const size_t cap = lds / (2 * elem);
    *longp = false;
    if (smooth5(L) && size_t(L) <= cap) return make_plan(L, lds, p);
    if (!smooth5(L) && size_t(next_smooth5(2 * L - 1)) <= cap) return make_plan(L, lds, p);
    *longp = true;
    return make_long_plan(L, lds, lp);
}

int check_sizes(const char* name, int M, int N)
{
    if (M < 2 || N < 2 || M % 2 || N % 2) {
        set_error(std::string(name) + ": sizes must be even and >= 2");
        return GSMI_E_ARG;
    }
    if (size_t(M) * size_t(N) > (size_t(1) << 31)) {
        set_error(std::string(name) + ": image larger than 2^31 elements");
        return GSMI_E_SIZE;
    }
    return GSMI_OK;
}

}  // namespace
}  // namespace gsmi

using namespace gsmi;

extern "C" {

int gsmi_cfft2_f64(const double* in, int M, int N, int inverse, int shift_in, int shift_out,
                   double* out, void* scratch, void* stream)
{
    if (!in || !out || !scratch) {
        set_error("gsmi_cfft2_f64: null buffer");
        return GSMI_E_ARG;
    }
    int rc = check_sizes("gsmi_cfft2_f64", M, N);
    if (rc) return rc;
    const size_t lds = lds_limit();
    LinePlan<double> px, py;
    LongPlan lpy;
    bool ly = false;
    if ((rc = make_plan(N, lds, &px)) || (rc = column_plan(M, lds, &py, &lpy, &ly))) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const auto* src = reinterpret_cast<const cplx<double>*>(in);
    auto* tmp = reinterpret_cast<cplx<double>*>(scratch);
    auto* dst = reinterpret_cast<cplx<double>*>(out);
    const int conj_io = inverse ? 0 : 1;
    const double scale = inverse ? 1. / (double(M) * double(N)) : 1.;
    // rows: input rotated in both axes (shift_in), output x rotated (shift_out)
    rc = launch_c2c(src, tmp, M, N, 1, shift_in ? M / 2 : 0, shift_in ? N / 2 : 0,
                    shift_out ? N / 2 : 0, conj_io, 1., px, lds, s);
    if (rc) return rc;
    // columns: output y rotated (shift_out)
    if (ly) return launch_columns_long(tmp, tmp, dst, N, lpy, 0, shift_out, conj_io, scale, lds, s);
    return launch_c2c(tmp, dst, N, 1, N, 0, 0, shift_out ? M / 2 : 0, conj_io, scale, py, lds, s);
}

int gsmi_rfft2_f64(const double* in, int M, int N, int shift_in, int shift_out, double* out,
                   void* scratch, void* stream)
{
    if (!in || !out || !scratch) {
        set_error("gsmi_rfft2_f64: null buffer");
        return GSMI_E_ARG;
    }
    int rc = check_sizes("gsmi_rfft2_f64", M, N);
    if (rc) return rc;
    const size_t lds = lds_limit();
    LinePlan<double> px, py;
    LongPlan lpy;
    bool ly = false;
    if ((rc = make_plan(N / 2, lds, &px)) || (rc = column_plan(M, lds, &py, &lpy, &ly)))
        return rc;
    const cplx<double>* twn = twiddles<double>(N);
    if (!twn) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    auto* tmp = reinterpret_cast<cplx<double>*>(scratch);
    const int W = N / 2 + 1;
    if (reg_len(N / 2) && aligned16(in) && aligned16(tmp)) {
        rc = launch_r2c_reg(N / 2, in, tmp, M, shift_in ? M / 2 : 0, shift_in ? N / 2 : 0, s);
    } else if (px.chirp && blue_len(N / 2)) {
        rc = launch_blue_rows(true, in, tmp, M, N, shift_in ? M / 2 : 0, shift_in ? N / 2 : 0, 1.,
                              s);
    } else {
        const int nl = lines_per_group(px.S, M, lds);
        const size_t bytes = 2 * size_t(nl) * px.S * sizeof(cplx<double>);
        if ((rc = allow_lds(reinterpret_cast<const void*>(r2c_rows<double>), bytes))) return rc;
        hipLaunchKernelGGL(r2c_rows<double>, dim3((M + nl - 1) / nl), dim3(kThreads), bytes, s,
                           in, tmp, M, N, nl, shift_in ? M / 2 : 0, shift_in ? N / 2 : 0, px,
                           twn);
        rc = hip_check(hipGetLastError(), "r2c_rows launch");
    }
    if (rc) return rc;
    if (ly)
        return launch_columns_long(tmp, tmp, reinterpret_cast<cplx<double>*>(out), W, lpy, 0,
                                   shift_out, 1, 1., lds, s);
    return launch_c2c(tmp, reinterpret_cast<cplx<double>*>(out), W, 1, W, 0, 0,
                      shift_out ? M / 2 : 0, 1, 1., py, lds, s);
}

int gsmi_irfft2_f64(const double* in, int M, int N, int shift_in, int shift_out, double* out,
                    void* scratch, void* stream)
{
    if (!in || !out || !scratch) {
        set_error("gsmi_irfft2_f64: null buffer");
        return GSMI_E_ARG;
    }
    int rc = check_sizes("gsmi_irfft2_f64", M, N);
    if (rc) return rc;
    const size_t lds = lds_limit();
    LinePlan<double> px, py;
    LongPlan lpy;
    bool ly = false;
    if ((rc = make_plan(N / 2, lds, &px)) || (rc = column_plan(M, lds, &py, &lpy, &ly)))
        return rc;
    const cplx<double>* twn = twiddles<double>(N);
    if (!twn) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    auto* tmp = reinterpret_cast<cplx<double>*>(scratch);
    const int W = N / 2 + 1;
    // columns first (inverse, input rows rotated by shift_in), unscaled; a four-step column
    // transform leaves its result in the second half of scratch
    if (ly) {
        rc = launch_columns_long(reinterpret_cast<const cplx<double>*>(in), tmp,
                                 tmp + size_t(M) * W, W, lpy, shift_in, 0, 0, 1., lds, s);
        tmp += size_t(M) * W;
    } else {
        rc = launch_c2c(reinterpret_cast<const cplx<double>*>(in), tmp, W, 1, W, 0,
                        shift_in ? M / 2 : 0, 0, 0, 1., py, lds, s);
    }
    if (rc) return rc;
    if (reg_len(N / 2) && aligned16(tmp) && aligned16(out))
        return launch_c2r_reg(N / 2, tmp, out, M, shift_out ? M / 2 : 0, shift_out ? N / 2 : 0,
                              1. / (double(M) * double(N)), s);
    if (px.chirp && blue_len(N / 2))
        return launch_blue_rows(false, tmp, out, M, N, shift_out ? M / 2 : 0,
                                shift_out ? N / 2 : 0, 1. / (double(M) * double(N)), s);
    const int nl = lines_per_group(px.S, M, lds);
    const size_t bytes = 2 * size_t(nl) * px.S * sizeof(cplx<double>);
    if ((rc = allow_lds(reinterpret_cast<const void*>(c2r_rows<double>), bytes))) return rc;
    hipLaunchKernelGGL(c2r_rows<double>, dim3((M + nl - 1) / nl), dim3(kThreads), bytes, s, tmp,
                       out, M, N, nl, shift_out ? M / 2 : 0, shift_out ? N / 2 : 0,
                       1. / (double(M) * double(N)), px, twn);
    return hip_check(hipGetLastError(), "c2r_rows launch");
}

int gsmi_fft_column_passes(int M)
{
    if (M < 2 || M % 2) {
        set_error("gsmi_fft_column_passes: M must be even and >= 2");
        return GSMI_E_ARG;
    }
    LinePlan<double> p;
    LongPlan lp;
    bool longp = false;
    const int rc = column_plan(M, lds_limit(), &p, &lp, &longp);
    if (rc) return rc;
    return longp ? 2 : 1;
}

int gsmi_hermitian_fill_f64(const double* half, int M, int N, int shift_out, double* out,
                            void* stream)
{
    if (!half || !out) {
        set_error("gsmi_hermitian_fill_f64: null buffer");
        return GSMI_E_ARG;
    }
    int rc = check_sizes("gsmi_hermitian_fill_f64", M, N);
    if (rc) return rc;
    const size_t total = size_t(M) * N;
    hipLaunchKernelGGL(hermitian_fill<double>, dim3(unsigned((total + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const cplx<double>*>(half), M, N, shift_out,
                       reinterpret_cast<cplx<double>*>(out));
    return hip_check(hipGetLastError(), "hermitian_fill launch");
}

}  // extern "C"
