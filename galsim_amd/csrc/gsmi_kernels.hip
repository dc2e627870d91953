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

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsmi.h"
#include "device_program.h"
#include "launch.h"

namespace gsmi {


// ---------------------------------------------------------------------------------------------
// Kernel 0: per-stamp separable axis factors X(kx), Y(ky) into the workspace (grid: nstamps)
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
axis_factors(gsmi_stamp* __restrict__ stamps, const gsmi_op* __restrict__ ops,
             const gsmi_leaf* __restrict__ leaves, cplx<T>* __restrict__ ws)
{
    gsmi_stamp* st = stamps + blockIdx.x;
    if (threadIdx.x == 0) st->status = 0;         // this launch's status starts clear
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

// The wrap: target cell (c, r) of the wrapped N-grid (c = 0..N/2, r = 0..N-1) is the sum of
// every k-image sample (kx, ky) congruent to it modulo N with |kx|, |ky| <= hk = Nk/2 (the stored
// sample, conjugated where it stands for kx < 0) of program x axis factors X(kx) Y(ky).  Every
// cell has the primary sample (c, kyp), kyp = r < N/2 ? r : r - N (Nk >= N); only the fringe has
// others: columns c >= a and rows |kyp| >= a, a = N - hk (every cell when a <= 0).  So the cells
// are filled in two passes: one primary sample per cell, straight-line, then the fringe cells,
// listed compactly (every column of the chunk x the fringe rows, then the fringe columns x the
// other rows), each adding its other samples.
//
// Columns kx = 0 and kx = N/2 are Hermitian along ky (F(-k) = conj F(k)), so their transforms
// along ky are real -- the real inverse transform along x reads only their real parts.  They
// travel as one complex column: column 0 holds cell(0, r) + i cell(N/2, r), and the column
// transform gives column 0 in its real part and column N/2 in its imaginary part.  The fold
// therefore transforms N/2 columns, c = 0..N/2-1 (column N/2's samples are gathered into
// column 0 by the fringe pass).

// Row of fringe-row index f (0 <= f < fringe_rows): kyp = a.. N/2-1, then -N/2 .. -a.
__device__ __forceinline__ int fringe_row(int f, int a, int N)
{
    if (a <= 0) return f;
    const int up = N / 2 - a;
    const int kyp = f < up ? a + f : -N / 2 + (f - up);
    return kyp < 0 ? kyp + N : kyp;
}

// Row of non-fringe-row index g (0 <= g < 2a - 1): kyp = -(a-1) .. a-1.
__device__ __forceinline__ int inner_row(int g, int a, int N)
{
    const int kyp = g - (a - 1);
    return kyp < 0 ? kyp + N : kyp;
}

// Fill cells of columns [c0, c0 + ncol) (within 0..N/2-1) into bufa (column-major in the padded
// transform layout: cell (cl, r) at cl * lds_line(N) + lds_pad(r)).  value(i, j) is the
// contribution of the stored sample (i, j) = (|kx|, kx < 0 ? -ky : ky), conjugated here when it
// stands for kx < 0.  Ends with a barrier.
template <typename T, typename V>
__device__ __forceinline__ void fold_cells_split(V&& value, int N, int Nk, int c0, int ncol,
                                                 cplx<T>* bufa)
{
    const int hk = Nk / 2;
    const int a = N - hk;
    const int LN = lds_line(N);
    const float inv_n = 1.f / float(N);
    for (int e = threadIdx.x; e < ncol * N; e += blockDim.x) {
        const int cl = fdiv(e, inv_n);
        const int r = e - cl * N;
        const int kyp = r < N / 2 ? r : r - N;
        bufa[cl * LN + lds_pad(r)] = value(c0 + cl, kyp);
    }
    __syncthreads();
    // every sample congruent to (c, r) except (skip_c, skip_ky)
    auto gather = [&](int c, int r, int skip_c, int skip_ky) -> cplx<T> {
        const int kx0 = c - N * fdiv(c + hk, inv_n);
        const int ky0 = r - N * fdiv(r + hk, inv_n);
        cplx<T> acc = {T(0), T(0)};
        for (int kx = kx0; kx <= hk; kx += N) {
            for (int ky = ky0; ky <= hk; ky += N) {
                if (kx == skip_c && ky == skip_ky) continue;
                const bool neg = kx < 0;
                const cplx<T> v = value(neg ? -kx : kx, neg ? -ky : ky);
                acc.x += v.x;
                acc.y += neg ? -v.y : v.y;
            }
        }
        return acc;
    };
    const int nfr = a <= 0 ? N : N - 2 * a + 1;
    const int cfirst = max(c0, a);
    const int ncf = (a <= 0) ? 0 : max(0, c0 + ncol - cfirst);
    const int ninner = 2 * a - 1;
    const int part1 = ncol * nfr;
    const int part2 = part1 + ncf * ninner;
    // column 0's rows outside the fringe rows still take column N/2's samples
    const int total = part2 + ((c0 == 0 && a > 0) ? ninner : 0);
    const float inv_nfr = 1.f / float(nfr);
    const float inv_inner = 1.f / float(max(ninner, 1));
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int cl, r;
        if (e < part1) {
            cl = fdiv(e, inv_nfr);
            r = fringe_row(e - cl * nfr, a, N);
        } else if (e < part2) {
            const int q = e - part1;
            const int k = fdiv(q, inv_inner);
            cl = cfirst - c0 + k;
            r = inner_row(q - k * ninner, a, N);
        } else {
            cl = 0;
            r = inner_row(e - part2, a, N);
        }
        const int c = c0 + cl;
        const int kyp = r < N / 2 ? r : r - N;
        // the primary sample is already in bufa; a cell of column 0 outside the fringe has no
        // other samples of its own, only column N/2's
        const bool own = (e < part2);
        cplx<T> acc = own ? gather(c, r, c, kyp) : cplx<T>{T(0), T(0)};
        if (c == 0) {
            const cplx<T> h = gather(N / 2, r, -1, 0);      // + i cell(N/2, r)
            acc.x -= h.y;
            acc.y += h.x;
        }
        cplx<T>& cell = bufa[cl * LN + lds_pad(r)];
        cell.x += acc.x;
        cell.y += acc.y;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void fold_cells(const gsmi_op* __restrict__ sops, int nops,
                                           const gsmi_leaf* __restrict__ sleaves,
                                           const Tables<T>& tb, const cplx<T>* __restrict__ X,
                                           const cplx<T>* __restrict__ Y, T dk, int N, int Nk,
                                           int c0, int ncol, cplx<T>* bufa)
{
    auto value = [&](int i, int j) -> cplx<T> {
        const cplx<T> v = eval_program<T>(sops, nops, sleaves, tb, dk, i, j, Nk);
        return cmul(v, cmul(X[i], Y[j]));
    };
    fold_cells_split<T>(value, N, Nk, c0, ncol, bufa);
}

// fold_cells for a stamp whose program has the compile-time shape SIG (leaf constants in lc).
template <typename T, int SIG>
__device__ __forceinline__ void fold_cells_static(const LeafC<T>* __restrict__ lc,
                                                  const LeafI (&li)[kProgLeaves],
                                                  const gsmi_leaf* __restrict__ sleaves,
                                                  const Tables<T>& tb,
                                                  const cplx<T>* __restrict__ X,
                                                  const cplx<T>* __restrict__ Y, T dk, int N,
                                                  int Nk, int c0, int ncol, cplx<T>* bufa)
{
    LeafC<T> lr[kProgLeaves];
#pragma unroll
    for (int l = 0; l < kStaticPrograms[SIG].nl; ++l) lr[l] = uniform_leaf(lc[l]);
    auto value = [&](int i, int j) -> cplx<T> {
        const T P = eval_static<T, SIG>(lr, li, sleaves, tb, T(i) * dk, T(j) * dk, i, j, Nk);
        const cplx<T> xy = cmul(X[i], Y[j]);
        return {P * xy.x, P * xy.y};
    };
    fold_cells_split<T>(value, N, Nk, c0, ncol, bufa);
}

// ---------------------------------------------------------------------------------------------
// Kernel 1: fused k-space evaluation + product/sinc/phase + Hermitian wrap + column inverse DFT
// grid: (ceil((N/2)/cpg), nstamps); dynamic LDS: 2 * cpg * lds_line(N) complex
// ---------------------------------------------------------------------------------------------

// Columns [first, first + count) of one stamp in chunks of at most cpg: each chunk's cells are
// evaluated into bufa, transformed along ky in LDS (twiddles twp, stride twstride) and the rows
// of the output window stored.  STATIC: the program has the compile-time shape SIG (leaf
// constants lc); otherwise the interpreter runs.  X holds the axis factors X(kx), kx in [0, hk],
// then Y(ky), ky in [-hk, hk].
// The interpreter's cell evaluation, out of line for the specialised kernels: there it only runs
// for stamps that do not match the compiled program, and inlined it would set the register
// budget of the whole kernel.
template <typename T>
__device__ __noinline__ void fold_cells_interp(const gsmi_op* __restrict__ sops, int nops,
                                               const gsmi_leaf* __restrict__ sleaves,
                                               Tables<T> tb, const cplx<T>* X, const cplx<T>* Y,
                                               T dk, int N, int Nk, int c0, int ncol,
                                               cplx<T>* bufa)
{
    fold_cells<T>(sops, nops, sleaves, tb, X, Y, dk, N, Nk, c0, ncol, bufa);
}

template <typename T, int SIG, bool STATIC>
__device__ __forceinline__ void fold_chunks(const gsmi_stamp* __restrict__ st, int first,
                                            int count, int cpg, const LeafC<T>* lc,
                                            const LeafI (&li)[kProgLeaves],
                                            const gsmi_op* __restrict__ sops, int nops,
                                            const gsmi_leaf* __restrict__ sleaves,
                                            const Tables<T>& tb, const cplx<T>* X,
                                            const cplx<T>* twp, int twstride,
                                            cplx<T>* __restrict__ scratch, cplx<T>* bufa,
                                            cplx<T>* bufb)
{
    const int N = st->N;
    const int Nk = st->Nk;
    const int hk = Nk / 2;
    const T dk = T(st->dk);
    cplx<T>* out = scratch + st->scratch_offset;
    const int ny = st->ny;
    const size_t cs = col_stride(st);
    const int y0 = st->y0;
    const float inv_n = 1.f / float(N);
    for (int c0 = first; c0 < first + count; c0 += cpg) {
        const int ncol = min(cpg, first + count - c0);
        if constexpr (STATIC)
            fold_cells_static<T, SIG>(lc, li, sleaves, tb, X, X + 2 * hk + 1, dk, N, Nk, c0,
                                      ncol, bufa);
        else if constexpr (SIG > 0)
            fold_cells_interp<T>(sops, nops, sleaves, tb, X, X + 2 * hk + 1, dk, N, Nk, c0, ncol,
                                 bufa);
        else
            fold_cells<T>(sops, nops, sleaves, tb, X, X + 2 * hk + 1, dk, N, Nk, c0, ncol, bufa);
        cplx<T>* res = fft_lds_c<T, true>(bufa, bufb, N, ncol, twp, twstride);
        // rows of the output window (position p <-> DFT index p mod N), column-major, columns
        // cs apart
        const int LN = lds_line(N);
        const float inv_ny = 1.f / float(ny);
        cplx<T>* oc = out + size_t(c0) * cs;
        for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
            const int cl = fdiv(e, inv_ny);
            const int row = e - cl * ny;
            const int p = y0 + row + N;               // >= 0: y0 >= -N/2
            const int y = p - N * fdiv(p, inv_n);
            const cplx<T> v = res[cl * LN + lds_pad(y)];
            if (c0 + cl == 0) {
                // the packed column: column 0 in the real part, column N/2 in the imaginary part
                oc[row] = {v.x, T(0)};
                out[size_t(N / 2) * cs + row] = {v.y, T(0)};
            } else {
                oc[cl * cs + row] = v;
            }
        }
        __syncthreads();                          // bufa / bufb are reused by the next chunk
    }
}

// One fold work item: columns [first, first + count) of stamp st, cpg columns at a time, in the
// workgroup's LDS of lds_elems complex values.  Set-up is done once per item: the program check
// and its leaf constants (shape SIG), and, when they fit after the two transform buffers, the
// stamp's axis factors and the N twiddles in LDS; otherwise a one-chunk item stages the axis
// factors in bufb (free until the column DFT), and the rest read them from global memory.
template <typename T, int SIG = 0>
__device__ __forceinline__ void fold_item(const gsmi_stamp* __restrict__ st, int first,
                                          int count, int cpg, const gsmi_op* __restrict__ ops,
                                          const gsmi_leaf* __restrict__ leaves,
                                          const Tables<T>& tb, const cplx<T>* __restrict__ tw,
                                          int tws, cplx<T>* __restrict__ scratch, cplx<T>* lds,
                                          size_t lds_elems)
{
    const int N = st->N;
    const gsmi_op* sops = ops + st->op_begin;
    const int nops = st->op_count;
    const gsmi_leaf* sleaves = leaves + st->leaf_begin;
    const int hk = st->Nk / 2;
    const cplx<T>* Xg = scratch + st->axis_offset;    // X[kx], kx in [0, hk]; then Y[ky]
    const int LN = lds_line(N);
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + cpg * LN;
    const int naxis = 3 * hk + 2;
    size_t used = size_t(2) * cpg * LN;
    LeafC<T>* lc = reinterpret_cast<LeafC<T>*>(lds + used);
    LeafI li[kProgLeaves];
    bool specialised = false;
    if constexpr (SIG > 0) {
        used += kProgLeaves * sizeof(LeafC<T>) / sizeof(cplx<T>);
        specialised = matches_static<SIG>(sops, nops, sleaves);
        if (specialised) {
#pragma unroll
            for (int l = 0; l < kProgLeaves; ++l)
                li[l] = (l < kStaticPrograms[SIG].nl) ? leaf_ints(sleaves[l])
                                                      : LeafI{0, 0, 0, -1, 0};
            if (int(threadIdx.x) < kStaticPrograms[SIG].nl)
                lc[threadIdx.x] = leaf_consts_t<T>(sleaves[threadIdx.x], tb);
        }
    }
    // FP32 programs with a Sersic leaf read its cubic coefficients from LDS: staged here, and a
    // stamp whose coefficients do not fit takes the interpreter
    // (the specialised path's coefficient pointer is the LDS address whether or not a stamp
    // uses it, so that its loads compile to LDS reads)
    Tables<T> tbs = tb;
    if constexpr (sizeof(T) == 4 && sersic_slot<SIG>() >= 0) {
        float4* cfl = reinterpret_cast<float4*>(lds + used);
        tbs.coef = cfl;
        if (specialised) {
            constexpr int ss = sersic_slot<SIG>();
            const int ncf = sleaves[ss].ntab - 1;
            const int cfo = sleaves[ss].cliprows;
            if (cfo < 0 || ncf < 1 || used + 2 * size_t(ncf) > lds_elems) {
                specialised = false;
            } else {
                const float4* src = tb.coef + cfo / 4;
                for (int e = threadIdx.x; e < ncf; e += blockDim.x) cfl[e] = src[e];
                li[ss].cf = 0;
                used += 2 * size_t(ncf);
            }
        }
    }
    cplx<T>* twl = lds + used;
    cplx<T>* axl = twl + N;
    if (used + N + naxis <= lds_elems) {
        for (int e = threadIdx.x; e < N; e += blockDim.x) twl[e] = tw[e * tws];
        for (int e = threadIdx.x; e < naxis; e += blockDim.x) axl[e] = Xg[e];
        __syncthreads();
        if constexpr (SIG > 0)
            if (specialised) {
                fold_chunks<T, SIG, true>(st, first, count, cpg, lc, li, sops, nops, sleaves, tbs,
                                          axl, twl, 1, scratch, bufa, bufb);
                return;
            }
        fold_chunks<T, SIG, false>(st, first, count, cpg, lc, li, sops, nops, sleaves, tb, axl,
                                   twl, 1, scratch, bufa, bufb);
        return;
    }
    if (count <= cpg && naxis <= cpg * LN) {
        for (int e = threadIdx.x; e < naxis; e += blockDim.x) bufb[e] = Xg[e];
        __syncthreads();
        if constexpr (SIG > 0)
            if (specialised) {
                fold_chunks<T, SIG, true>(st, first, count, cpg, lc, li, sops, nops, sleaves, tbs,
                                          bufb, tw, tws, scratch, bufa, bufb);
                return;
            }
        fold_chunks<T, SIG, false>(st, first, count, cpg, lc, li, sops, nops, sleaves, tb, bufb,
                                   tw, tws, scratch, bufa, bufb);
        return;
    }
    __syncthreads();
    if constexpr (SIG > 0)
        if (specialised) {
            fold_chunks<T, SIG, true>(st, first, count, cpg, lc, li, sops, nops, sleaves, tbs, Xg,
                                      tw, tws, scratch, bufa, bufb);
            return;
        }
    fold_chunks<T, SIG, false>(st, first, count, cpg, lc, li, sops, nops, sleaves, tb, Xg, tw,
                               tws, scratch, bufa, bufb);
}

// One stamp per blockIdx.y, column group blockIdx.x (all stamps of the launch share N).
template <typename T>
__device__ __forceinline__ void fold_colfft_body(const gsmi_stamp* __restrict__ stamps,
                                                 const gsmi_op* __restrict__ ops,
                                                 const gsmi_leaf* __restrict__ leaves,
                                                 const Tables<T>& tb,
                                                 const cplx<T>* __restrict__ tw,
                                                 cplx<T>* __restrict__ scratch, int cpg)
{
    extern __shared__ unsigned char smem_raw[];
    const gsmi_stamp* st = stamps + blockIdx.y;
    const int N = st->N;
    const int c0 = blockIdx.x * cpg;
    const int ncol = min(cpg, N / 2 - c0);
    if (ncol <= 0) return;
    fold_item<T>(st, c0, ncol, cpg, ops, leaves, tb, tw, 1, scratch,
                 reinterpret_cast<cplx<T>*>(smem_raw), size_t(2) * cpg * lds_line(N));
}

// LDS a work-list fold launch needs beyond its two transform buffers (the decoded leaf constants
// of a specialised program).
template <typename T>
constexpr size_t kFoldLdsExtra = kProgLeaves * sizeof(LeafC<T>);

// Work-list form: every FFT size in one launch.  Workgroup b takes items b, b + grid, ...
// (persistent when the grid is smaller than the list); an item (stamp, first column, columns,
// columns per chunk) may span several chunks, which share its set-up.  Items are validated
// against the stamp records and the launch's LDS so a malformed list skips items instead of
// overrunning memory.
template <typename T, int SIG>
__device__ __forceinline__ void fold_list_body(const gsmi_stamp* __restrict__ stamps, int nstamps,
                                               const gsmi_work* __restrict__ items, int nitems,
                                               int lds_bytes, const gsmi_op* __restrict__ ops,
                                               const gsmi_leaf* __restrict__ leaves,
                                               const Tables<T>& tb, TwiddleSet<T> tws,
                                               cplx<T>* __restrict__ scratch)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const gsmi_work it = items[w];
        if (it.stamp < 0 || it.stamp >= nstamps) continue;
        const gsmi_stamp* st = stamps + it.stamp;
        const int N = st->N;
        if (it.count <= 0 || it.first < 0 || it.first + it.count > N / 2 ||
            it.per_group <= 0 ||
            size_t(2) * it.per_group * lds_line(N) * sizeof(cplx<T>) +
                    (SIG > 0 ? kFoldLdsExtra<T> : 0) >
                size_t(lds_bytes))
            continue;
        int stride;
        const cplx<T>* tw = tws.pick(N, &stride);
        fold_item<T, SIG>(st, it.first, it.count, it.per_group, ops, leaves, tb, tw, stride,
                          scratch, lds, size_t(lds_bytes) / sizeof(cplx<T>));
        __syncthreads();
    }
}

// The FP32 kernel is held to 96 VGPRs (5 waves per SIMD): its sample loop is latency bound and
// runs faster with the extra wave than without the few registers it gives up.
__global__ void __launch_bounds__(kThreads, 5)
fold_colfft_f32(const gsmi_stamp* __restrict__ stamps, const gsmi_op* __restrict__ ops,
                const gsmi_leaf* __restrict__ leaves, Tables<float> tb,
                const cplx<float>* __restrict__ tw, cplx<float>* __restrict__ scratch, int cpg)
{
    fold_colfft_body<float>(stamps, ops, leaves, tb, tw, scratch, cpg);
}

__global__ void __launch_bounds__(kThreads)
fold_colfft_f64(const gsmi_stamp* __restrict__ stamps, const gsmi_op* __restrict__ ops,
                const gsmi_leaf* __restrict__ leaves, Tables<double> tb,
                const cplx<double>* __restrict__ tw, cplx<double>* __restrict__ scratch, int cpg)
{
    fold_colfft_body<double>(stamps, ops, leaves, tb, tw, scratch, cpg);
}

template <int SIG>
__global__ void __launch_bounds__(kThreads, 5)
fold_list_f32(const gsmi_stamp* __restrict__ stamps, int nstamps,
              const gsmi_work* __restrict__ items, int nitems, int lds_bytes,
              const gsmi_op* __restrict__ ops, const gsmi_leaf* __restrict__ leaves,
              Tables<float> tb, TwiddleSet<float> tws, cplx<float>* __restrict__ scratch)
{
    fold_list_body<float, SIG>(stamps, nstamps, items, nitems, lds_bytes, ops, leaves, tb, tws,
                               scratch);
}

template <int SIG>
__global__ void __launch_bounds__(kThreads)
fold_list_f64(const gsmi_stamp* __restrict__ stamps, int nstamps,
              const gsmi_work* __restrict__ items, int nitems, int lds_bytes,
              const gsmi_op* __restrict__ ops, const gsmi_leaf* __restrict__ leaves,
              Tables<double> tb, TwiddleSet<double> tws, cplx<double>* __restrict__ scratch)
{
    fold_list_body<double, SIG>(stamps, nstamps, items, nitems, lds_bytes, ops, leaves, tb, tws,
                                scratch);
}

template <typename T>
struct FoldKernel;
template <>
struct FoldKernel<float> {
    static constexpr auto fn = fold_colfft_f32;
    template <int SIG>
    static constexpr auto list = fold_list_f32<SIG>;
};
template <>
struct FoldKernel<double> {
    static constexpr auto fn = fold_colfft_f64;
    template <int SIG>
    static constexpr auto list = fold_list_f64<SIG>;
};

// ---------------------------------------------------------------------------------------------
// Kernel 2: real inverse DFT along x of the column-transformed rows, crop and scale.
// Row z holds Z(c), c = 0..N/2; the real sequence x(q) = sum_c Z(c) e^{2 pi i c q / N} (Hermitian
// extension) is formed as v = IDFT_{N/2}(E + i O), x(2m) = Re v(m), x(2m+1) = Im v(m), with
// E(c) = Z(c) + conj(Z(N/2-c)) and O(c) = (Z(c) - conj(Z(N/2-c))) e^{2 pi i c / N}.
// ---------------------------------------------------------------------------------------------
// Row rl, column c of the input is in[rl * rs + c * cs]: column-major scratch of the drawing
// kernels (rs = 1, cs = the stamp's rows; threads take consecutive rows of a column, so a wave
// reads whole cache lines), or row-major (cs = 1; threads take consecutive columns).
template <typename T>
__device__ __forceinline__ void row_c2r_body(cplx<T>* lds, int rpg, const cplx<T>* __restrict__ in,
                                             int rs, int cs, int nrow, int N, int x0, int nx,
                                             const cplx<T>* __restrict__ tw,
                                             T* __restrict__ out, int tws, bool& bad)
{
    const int n2 = N / 2;
    const int L2 = lds_line(n2);
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + rpg * L2;
    const float inv_n2 = 1.f / float(n2);
    const float inv_nrow = 1.f / float(nrow);
    for (int e = threadIdx.x; e < nrow * n2; e += blockDim.x) {
        int rl, c;
        if (rs == 1) {
            c = fdiv(e, inv_nrow);
            rl = e - c * nrow;
        } else {
            rl = fdiv(e, inv_n2);
            c = e - rl * n2;
        }
        const cplx<T>* z = in + size_t(rl) * rs;
        cplx<T> zc = z[size_t(c) * cs];
        cplx<T> zm = z[size_t(n2 - c) * cs];
        if (c == 0) { zc.y = T(0); zm.y = T(0); }    // DC and Nyquist are real
        cplx<T> E = cadd(zc, cconj(zm));
        cplx<T> O = cmul(csub(zc, cconj(zm)), tw[c * tws]);
        bufa[rl * L2 + lds_pad(c)] = {E.x - O.y, E.y + O.x};  // E + i O
    }
    __syncthreads();
    cplx<T>* res = fft_lds_c<T, true>(bufa, bufb, n2, nrow, tw, 2 * tws);
    const T norm = T(1) / (T(N) * T(N));
    const float inv_nx = 1.f / float(nx);
    const float inv_n = 1.f / float(N);
    for (int e = threadIdx.x; e < nrow * nx; e += blockDim.x) {
        const int rl = fdiv(e, inv_nx);
        const int col = e - rl * nx;
        const int p = x0 + col + N;                   // >= 0: x0 >= -N/2
        const int q = p - N * fdiv(p, inv_n);
        const cplx<T> v = res[rl * L2 + lds_pad(q >> 1)];
        const T px = (q & 1) ? v.y : v.x;
        out[size_t(rl) * nx + col] = px * norm;
        bad |= !isfinite(px);
    }
}

// ---------------------------------------------------------------------------------------------
// The row transform with N a compile-time constant (the common FFT sizes).  The half-length
// transform of M = N/2 points runs as Stockham radix passes (3 first, then 8, 4, 2) over rows of
// the chunk held in LDS lines M + 1 apart (an odd stride: consecutive rows fall on different
// banks).  The first pass forms E + iO straight from the column-major scratch into registers
// (threads take consecutive rows of a column: whole cache lines), the middle passes go LDS to
// LDS (threads again take consecutive rows: no bank conflicts), and the last pass writes the
// cropped, 1/N^2-scaled pixels straight to the output (threads take consecutive outputs of a
// row: contiguous stores).  twl: exp(2 pi i k / N), k < N, in LDS.
// ---------------------------------------------------------------------------------------------
constexpr int row_radix(int rem)
{
    return (rem % 3 == 0) ? 3 : (rem % 8 == 0) ? 8 : (rem % 4 == 0) ? 4 : 2;
}

template <int R, typename T>
__device__ __forceinline__ void bfly(cplx<T>* v)
{
    if constexpr (R == 2) bfly2(v);
    if constexpr (R == 3) bfly3(v);
    if constexpr (R == 4) bfly4(v);
    if constexpr (R == 8) bfly8(v);
}

// The first pass reads each column pair once: the inputs of output c, X[c] and X[N/2 - c], give
// both v(c) = E + iO and v(N/2 - c) = conj(E) + i conj(O) (E(N/2 - c) = conj E(c) and, as
// e^{2 pi i (N/2 - c) / N} = -conj e^{2 pi i c / N}, O(N/2 - c) = conj O(c)).  Butterfly b
// (columns b + r NB) therefore runs with its partner NB - b (columns N/2 - b - r NB), whose
// input r is this one's R - 1 - r; b = 0 and b = NB / 2 are their own partners.
template <typename T, int N, int R>
__device__ __forceinline__ void rows_first(const cplx<T>* __restrict__ in, int cs, int nrow,
                                           float inv_nrow, const cplx<T>* twl, cplx<T>* dst)
{
    constexpr int M = N / 2, L = M + 1, NB = M / R, NP = NB / 2 + 1;
    const int total = nrow * NP;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int b = fdiv(t, inv_nrow);
        const int rl = t - b * nrow;
        const int b2 = NB - b;
        const bool pair = b != 0 && b2 != b;
        const cplx<T>* z = in + rl;
        cplx<T> v[R], vp[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int c = b + r * NB;
            cplx<T> zc = z[size_t(c) * cs];
            cplx<T> zm = z[size_t(M - c) * cs];
            if (r == 0 && b == 0) { zc.y = T(0); zm.y = T(0); }   // DC and Nyquist are real
            const cplx<T> E = cadd(zc, cconj(zm));
            const cplx<T> O = cmul(csub(zc, cconj(zm)), twl[c]);
            v[r] = {E.x - O.y, E.y + O.x};
            vp[R - 1 - r] = {E.x + O.y, -E.y + O.x};             // conj(E) + i conj(O)
        }
        bfly<R>(v);
        cplx<T>* d = dst + rl * L + b * R;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = v[r];
        if (pair) {
            bfly<R>(vp);
            cplx<T>* d2 = dst + rl * L + b2 * R;
#pragma unroll
            for (int r = 0; r < R; ++r) d2[r] = vp[r];
        }
    }
}

template <typename T, int N, int R, int NS>
__device__ __forceinline__ void rows_mid(const cplx<T>* src, cplx<T>* dst, int nrow,
                                         float inv_nrow, const cplx<T>* twl)
{
    constexpr int M = N / 2, L = M + 1, NB = M / R;
    constexpr int TSTEP = 2 * (M / (NS * R));
    const int total = nrow * NB;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int b = fdiv(t, inv_nrow);
        const int rl = t - b * nrow;
        const int tq = b / NS;
        const int jj = b - tq * NS;
        const cplx<T>* sp = src + rl * L + b;
        cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = sp[r * NB];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twl[jj * r * TSTEP]);
        bfly<R>(v);
        cplx<T>* d = dst + rl * L + tq * (NS * R) + jj;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * NS] = v[r];
    }
}

template <typename T, int N, int R, int NS>
__device__ __forceinline__ void rows_last(const cplx<T>* src, int nrow, const cplx<T>* twl,
                                          int x0, int nx, T* __restrict__ out, bool& bad)
{
    constexpr int M = N / 2, L = M + 1, NB = M / R;          // NB == NS
    const T norm = T(1) / (T(N) * T(N));
    const int total = nrow * NB;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int rl = t / NB;
        const int jj = t - rl * NB;
        const cplx<T>* sp = src + rl * L + jj;
        cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = sp[r * NB];
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twl[2 * jj * r]);
        }
        bfly<R>(v);
        // a non-finite value anywhere in the row's spectrum reaches every output of its
        // transform (each output sums every input with a unit-modulus weight, and no add or
        // multiply turns inf or NaN finite), so one output per thread stands for the row
        bad |= !isfinite(v[0].x + v[0].y);
        T* o = out + size_t(rl) * nx;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // output m = jj + r NS holds pixels q = 2m (real part) and 2m + 1 (imaginary part);
            // pixel q is column (q - x0) mod N of the window
            const int q = 2 * (jj + r * NS);
            int col = q - x0;
            col = col >= N ? col - N : (col < 0 ? col + N : col);
            int col1 = col + 1;
            col1 = col1 >= N ? col1 - N : col1;
            if (col < nx) o[col] = v[r].x * norm;
            if (col1 < nx) o[col1] = v[r].y * norm;
        }
    }
}

template <typename T, int N, int NS>
__device__ __forceinline__ void rows_rest(cplx<T>* src, cplx<T>* dst, int nrow, float inv_nrow,
                                          const cplx<T>* twl, int x0, int nx, T* out, bool& bad)
{
    constexpr int rem = (N / 2) / NS;
    constexpr int R = row_radix(rem);
    if constexpr (rem == R) {
        rows_last<T, N, R, NS>(src, nrow, twl, x0, nx, out, bad);
    } else {
        rows_mid<T, N, R, NS>(src, dst, nrow, inv_nrow, twl);
        __syncthreads();
        rows_rest<T, N, NS * R>(dst, src, nrow, inv_nrow, twl, x0, nx, out, bad);
    }
}

// nrow (<= rpg) rows of one stamp: in = the stamp's scratch + its first row (column stride cs),
// out = its first output row; LDS: 2 rpg (N/2 + 1) complex values.
template <typename T, int N>
__device__ __forceinline__ void rows_fixed(cplx<T>* lds, int rpg, const cplx<T>* __restrict__ in,
                                           int cs, int nrow, int x0, int nx, const cplx<T>* twl,
                                           T* __restrict__ out, bool& bad)
{
    constexpr int M = N / 2;
    constexpr int R1 = row_radix(M);
    cplx<T>* bufa = lds;
    cplx<T>* bufb = lds + rpg * (M + 1);
    const float inv_nrow = 1.f / float(nrow);
    rows_first<T, N, R1>(in, cs, nrow, inv_nrow, twl, bufa);
    __syncthreads();
    rows_rest<T, N, R1>(bufa, bufb, nrow, inv_nrow, twl, x0, nx, out, bad);
}

// rows_fixed for the compiled sizes; false (nothing done) for any other N.
template <typename T>
__device__ __forceinline__ bool rows_fixed_any(int N, cplx<T>* lds, int rpg,
                                               const cplx<T>* __restrict__ in, int cs, int nrow,
                                               int x0, int nx, const cplx<T>* twl,
                                               T* __restrict__ out, bool& bad)
{
    switch (N) {
#define GSMI_ROWS_FIXED(NN)                                                                     \
      case NN: rows_fixed<T, NN>(lds, rpg, in, cs, nrow, x0, nx, twl, out, bad); return true;
      GSMI_ROWS_FIXED(64)
      GSMI_ROWS_FIXED(96)
      GSMI_ROWS_FIXED(128)
      GSMI_ROWS_FIXED(192)
      GSMI_ROWS_FIXED(256)
      GSMI_ROWS_FIXED(384)
      GSMI_ROWS_FIXED(512)
      GSMI_ROWS_FIXED(768)
      GSMI_ROWS_FIXED(1024)
      GSMI_ROWS_FIXED(1536)
      GSMI_ROWS_FIXED(2048)
      GSMI_ROWS_FIXED(3072)
#undef GSMI_ROWS_FIXED
      default: return false;
    }
}

// grid: (ceil(ny/rpg), nstamps); dynamic LDS: 2 * rpg * N/2 complex
// A workgroup's verdict on its pixels: any non-finite one sets the stamp's status bit.
__device__ __forceinline__ void flag_nonfinite(gsmi_stamp* st, bool bad)
{
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&st->status, GSMI_STAMP_NONFINITE);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
row_c2r(gsmi_stamp* __restrict__ stamps, const cplx<T>* __restrict__ scratch,
        const cplx<T>* __restrict__ tw, T* __restrict__ out, int rpg)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    gsmi_stamp* st = stamps + blockIdx.y;
    const int N = st->N;
    const int r0 = blockIdx.x * rpg;
    const int nrow = min(rpg, st->ny - r0);
    if (nrow <= 0) return;
    bool bad = false;
    row_c2r_body<T>(lds, rpg, scratch + st->scratch_offset + r0, 1, col_stride(st), nrow, N,
                    st->x0, st->nx, tw, out + st->out_offset + size_t(r0) * st->nx, 1, bad);
    flag_nonfinite(st, bad);
}

// Work-list form of row_c2r: items (stamp, first row, rows, rows per chunk).  An item may span
// several chunks of per_group rows; the N twiddles are staged in LDS once per item when the
// launch's LDS has room for them after the two transform buffers.
template <typename T>
__global__ void __launch_bounds__(kThreads)
rows_list(gsmi_stamp* __restrict__ stamps, int nstamps,
          const gsmi_work* __restrict__ items, int nitems, int lds_bytes,
          const cplx<T>* __restrict__ scratch, TwiddleSet<T> tws, T* __restrict__ out)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const gsmi_work it = items[w];
        if (it.stamp < 0 || it.stamp >= nstamps) continue;
        gsmi_stamp* st = stamps + it.stamp;
        const int N = st->N;
        const size_t bufs = size_t(2) * it.per_group * lds_line(N / 2);
        if (it.count <= 0 || it.first < 0 || it.first + it.count > st->ny ||
            it.per_group <= 0 || bufs * sizeof(cplx<T>) > size_t(lds_bytes))
            continue;
        int stride;
        const cplx<T>* tw = tws.pick(N, &stride);
        const int end = it.first + it.count;
        bool bad = false;
        if ((bufs + N) * sizeof(cplx<T>) <= size_t(lds_bytes)) {
            cplx<T>* twl = lds + bufs;
            for (int e = threadIdx.x; e < N; e += blockDim.x) twl[e] = tw[e * stride];
            __syncthreads();
            for (int r0 = it.first; r0 < end; r0 += it.per_group) {
                const int nrow = min(it.per_group, end - r0);
                const cplx<T>* in = scratch + st->scratch_offset + r0;
                T* o = out + st->out_offset + size_t(r0) * st->nx;
                if (!rows_fixed_any<T>(N, lds, it.per_group, in, col_stride(st), nrow, st->x0,
                                       st->nx, twl, o, bad))
                    row_c2r_body<T>(lds, it.per_group, in, 1, col_stride(st), nrow, N, st->x0,
                                    st->nx, twl, o, 1, bad);
                __syncthreads();
            }
        } else {
            for (int r0 = it.first; r0 < end; r0 += it.per_group) {
                row_c2r_body<T>(lds, it.per_group, scratch + st->scratch_offset + r0, 1,
                                col_stride(st), min(it.per_group, end - r0), N, st->x0, st->nx, tw,
                                out + st->out_offset + size_t(r0) * st->nx, stride, bad);
                __syncthreads();
            }
        }
        flag_nonfinite(st, bad);
    }
}

// full N x N real image from N column-transformed rows (irfft entry point)
__global__ void __launch_bounds__(kThreads)
irfft_rows_f64(const cplx<double>* __restrict__ scratch, int N, const cplx<double>* __restrict__ tw,
               int tws, double* __restrict__ out)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
    const int rpg = rows_per_group(N);
    const int r0 = blockIdx.x * rpg;
    const int nrow = min(rpg, N - r0);
    if (nrow <= 0) return;
    const int W = N / 2 + 1;
    bool bad = false;
    row_c2r_body<double>(lds, rpg, scratch + size_t(r0) * W, W, 1, nrow, N, -N / 2, N, tw,
                         out + size_t(r0) * N, tws, bad);
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

// ApplyKImagePhases on a row-major complex128 k-image: pixel (x, y) = (xmin + i, ymin + j) at
// k = (x, y) dk is multiplied by fs exp(-i (kx' cenx + ky' ceny)), (kx', ky') = (A kx + C ky,
// B kx + D ky) for the Jacobian (A, B, C, D), the identity when none is given.
__global__ void __launch_bounds__(kThreads)
apply_kimage_phases_f64(cplx<double>* __restrict__ img, int xmin, int ymin, int nx, int ny,
                        double dk, double A, double B, double C, double D, double cenx,
                        double ceny, double fs)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny) return;
    const int j = idx / nx;
    const int i = idx - j * nx;
    const double kx = double(xmin + i) * dk;
    const double ky = double(ymin + j) * dk;
    const double arg = (A * kx + C * ky) * cenx + (B * kx + D * ky) * ceny;
    double sn, cs;
    sincos(arg, &sn, &cs);
    img[idx] = cmul(img[idx], cplx<double>{fs * cs, -fs * sn});
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
                 int tws, cplx<double>* __restrict__ scratch)
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
    cplx<double>* res = fft_lds<double>(bufa, bufb, N, ncol, tw, tws);
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
}  // namespace

template <>
const cplx<double>* twiddles<double>(int N) { return twd().get(N); }
template <>
const cplx<float>* twiddles<float>(int N) { return twf().get(N); }

// 2^k or 3*2^k with both in-LDS transforms within `lds` bytes per workgroup (64 KiB: N <= 2048
// in FP64, N <= 4096 in FP32; 160 KiB: N <= 4096 in FP64, N <= 8192 in FP32).
bool fft_size_ok(int N, size_t elem, size_t lds, bool padded)
{
    if (N < 4 || N % 2) return false;
    int m = N;
    while (m % 2 == 0) m /= 2;
    if (m != 1 && m != 3) return false;
    return 2 * size_t(cols_per_group(N)) * (padded ? lds_line(N) : N) * elem <= lds;
}

size_t lds_limit()
{
    // queried once per device (the per-stamp drawing path launches several kernels per stamp)
    static std::atomic<int> cached[64] = {};
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (dev >= 0 && dev < 64 && cached[dev].load(std::memory_order_relaxed) > 0)
        return size_t(cached[dev].load(std::memory_order_relaxed));
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v < 65536)
        v = 65536;
    if (dev >= 0 && dev < 64) cached[dev].store(v, std::memory_order_relaxed);
    return size_t(v);
}

int allow_lds(const void* kernel, size_t bytes)
{
    if (bytes <= 65536) return GSMI_OK;
    return hip_check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         int(bytes)),
                     "dynamic LDS attribute");
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
int launch_axis_factors(const gsmi_stamp* stamps, int nstamps, const gsmi_op* ops,
                        const gsmi_leaf* leaves, void* scratch, void* stream)
{
    // the kernel also clears each record's status word (the only field a kernel writes)
    hipLaunchKernelGGL(axis_factors<T>, dim3(nstamps), dim3(kThreads), 0,
                       reinterpret_cast<hipStream_t>(stream), const_cast<gsmi_stamp*>(stamps),
                       ops, leaves, reinterpret_cast<cplx<T>*>(scratch));
    return hip_check(hipGetLastError(), "axis_factors launch");
}
template int launch_axis_factors<double>(const gsmi_stamp*, int, const gsmi_op*,
                                         const gsmi_leaf*, void*, void*);
template int launch_axis_factors<float>(const gsmi_stamp*, int, const gsmi_op*,
                                        const gsmi_leaf*, void*, void*);

namespace {
// Columns (rows) per workgroup of a per-size launch over `nstamps` stamps of `lines` columns
// (rows) each: the LDS-filling tile `full`, halved while the launch would have fewer than
// 1024 workgroups, so that a launch of one or a few stamps spreads over the 256 CUs instead of
// running on a handful of them.  Any tile from 1 to `full` is valid for both kernels.
int small_launch_tile(int full, int lines, int nstamps)
{
    int t = full;
    while (t > 1 && size_t(nstamps) * size_t((lines + t - 1) / t) < 1024) t /= 2;
    return t;
}

template <typename T>
int fold_impl(const gsmi_stamp* stamps, int nstamps, int N, const gsmi_op* ops,
              const gsmi_leaf* leaves, Tables<T> tb, void* scratch, void* stream)
{
    if (!fft_size_ok(N, sizeof(cplx<T>), lds_limit(), true)) {
        set_error("unsupported FFT size " + std::to_string(N) +
                  " (need 2^k or 3*2^k, <= 4096 in FP64, <= 8192 in FP32)");
        return GSMI_E_SIZE;
    }
    if (nstamps <= 0) return GSMI_OK;
    const cplx<T>* tw = twiddles<T>(N);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    const int cpg = small_launch_tile(cols_per_group(N), N / 2, nstamps);
    int rc = launch_axis_factors<T>(stamps, nstamps, ops, leaves, scratch, stream);
    if (rc) return rc;
    dim3 g1((N / 2 + cpg - 1) / cpg, nstamps);
    size_t lds1 = 2 * size_t(cpg) * lds_line(N) * sizeof(cplx<T>);
    if ((rc = allow_lds(reinterpret_cast<const void*>(FoldKernel<T>::fn), lds1))) return rc;
    hipLaunchKernelGGL(FoldKernel<T>::fn, g1, dim3(kThreads), lds1,
                       reinterpret_cast<hipStream_t>(stream), stamps, ops, leaves, tb, tw,
                       reinterpret_cast<cplx<T>*>(scratch), cpg);
    return hip_check(hipGetLastError(), "fold_colfft launch");
}

template <typename T>
int rows_impl(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, const void* scratch,
              T* out, void* stream)
{
    if (!fft_size_ok(N, sizeof(cplx<T>), lds_limit(), true)) {
        set_error("unsupported FFT size " + std::to_string(N));
        return GSMI_E_SIZE;
    }
    if (nstamps <= 0) return GSMI_OK;
    const cplx<T>* tw = twiddles<T>(N);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    const int rpg = small_launch_tile(rows_per_group(N), max_ny, nstamps);
    dim3 g2((max_ny + rpg - 1) / rpg, nstamps);
    size_t lds2 = 2 * size_t(rpg) * lds_line(N / 2) * sizeof(cplx<T>);
    int rc = allow_lds(reinterpret_cast<const void*>(row_c2r<T>), lds2);
    if (rc) return rc;
    hipLaunchKernelGGL(row_c2r<T>, g2, dim3(kThreads), lds2, reinterpret_cast<hipStream_t>(stream),
                       const_cast<gsmi_stamp*>(stamps), reinterpret_cast<const cplx<T>*>(scratch),
                       tw, out, rpg);
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
template <typename T>
int sched_check(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sc)
{
    if (!stamps || !sc || nstamps < 0 || sc->nfold < 0 || sc->nrows < 0 ||
        (sc->nfold > 0 && (!sc->fold || sc->fold_grid <= 0)) ||
        (sc->nrows > 0 && (!sc->rows || sc->rows_grid <= 0)) || sc->fold_lds < 0 ||
        sc->rows_lds < 0 || size_t(sc->fold_lds) > lds_limit() ||
        size_t(sc->rows_lds) > lds_limit()) {
        set_error("bad schedule");
        return GSMI_E_ARG;
    }
    return GSMI_OK;
}

template <typename T>
int fold_list_impl(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sc,
                   const gsmi_op* ops, const gsmi_leaf* leaves, Tables<T> tb, void* scratch,
                   void* stream)
{
    int rc = sched_check<T>(stamps, nstamps, sc);
    if (rc) return rc;
    if (nstamps == 0) return GSMI_OK;
    TwiddleSet<T> tws = twiddle_set<T>();
    if (!tws.p2 || !tws.p3) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    rc = launch_axis_factors<T>(stamps, nstamps, ops, leaves, scratch, stream);
    if (rc || sc->nfold == 0) return rc;
    const dim3 grid(min(sc->fold_grid, sc->nfold));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    cplx<T>* scr = reinterpret_cast<cplx<T>*>(scratch);
    switch (sc->program) {
#define GSMI_FOLD_LIST(SIG)                                                                      \
      case SIG:                                                                                  \
        if ((rc = allow_lds(reinterpret_cast<const void*>(FoldKernel<T>::template list<SIG>),    \
                            sc->fold_lds)))                                                      \
            return rc;                                                                           \
        hipLaunchKernelGGL(FoldKernel<T>::template list<SIG>, grid, dim3(kThreads), sc->fold_lds, \
                           s, stamps, nstamps, sc->fold, sc->nfold, sc->fold_lds, ops, leaves,  \
                           tb, tws, scr);                                                       \
        break;
      GSMI_FOLD_LIST(0)
      GSMI_FOLD_LIST(1)
      GSMI_FOLD_LIST(2)
      GSMI_FOLD_LIST(3)
      GSMI_FOLD_LIST(4)
#undef GSMI_FOLD_LIST
      default:
        set_error("unknown program shape in schedule");
        return GSMI_E_ARG;
    }
    return hip_check(hipGetLastError(), "fold list launch");
}

template <typename T>
int rows_list_impl(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sc,
                   const void* scratch, T* out, void* stream)
{
    int rc = sched_check<T>(stamps, nstamps, sc);
    if (rc) return rc;
    if (nstamps == 0 || sc->nrows == 0) return GSMI_OK;
    TwiddleSet<T> tws = twiddle_set<T>();
    if (!tws.p2 || !tws.p3) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    if ((rc = allow_lds(reinterpret_cast<const void*>(rows_list<T>), sc->rows_lds))) return rc;
    hipLaunchKernelGGL(rows_list<T>, dim3(min(sc->rows_grid, sc->nrows)), dim3(kThreads),
                       sc->rows_lds, reinterpret_cast<hipStream_t>(stream),
                       const_cast<gsmi_stamp*>(stamps), nstamps,
                       sc->rows, sc->nrows, sc->rows_lds,
                       reinterpret_cast<const cplx<T>*>(scratch), tws, out);
    return hip_check(hipGetLastError(), "rows list launch");
}
}  // namespace

constexpr int kNoAllocMaxN = GSMI_NOALLOC_MAX_N;

// Column then row pass of irfft(shift_in, shift_out) through `scratch` (N (N/2+1) complex128),
// with twiddles tw[m * tws] = exp(2 pi i m / N).
int irfft_launch(const double* kin, int N, double* out, cplx<double>* scratch,
                 const cplx<double>* tw, int tws, void* stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int cpg = cols_per_group(N);
    size_t lds1 = 2 * size_t(cpg) * N * sizeof(cplx<double>);
    hipLaunchKernelGGL(irfft_colfft_f64, dim3((N / 2 + cpg) / cpg), dim3(kThreads), lds1, s,
                       reinterpret_cast<const cplx<double>*>(kin), N, tw, tws, scratch);
    int rc = hip_check(hipGetLastError(), "irfft columns launch");
    if (rc) return rc;
    const int rpg = rows_per_group(N);
    size_t lds2 = 2 * size_t(rpg) * lds_line(N / 2) * sizeof(cplx<double>);
    hipLaunchKernelGGL(irfft_rows_f64, dim3((N + rpg - 1) / rpg), dim3(kThreads), lds2, s,
                       scratch, N, tw, tws, out);
    return hip_check(hipGetLastError(), "irfft rows launch");
}

// ---------------------------------------------------------------------------------------------
// Static scratch of gsmi_irfft_f64_noalloc: the N (N/2+1) complex128 column-transformed rows for
// N <= kNoAllocMaxN, then the two twiddle tables exp(2 pi i m / 4096), m < 4096, and
// exp(2 pi i m / 3072), m < 3072, from which every supported N takes every (4096/N)-th or
// (3072/N)-th entry.  One copy per device, held until the process ends; every call on a device
// shares it, so calls that overlap on one device race.  Host code never passes the array's name
// to a kernel or copy: it uses the device address hipGetSymbolAddress returns for it.
// ---------------------------------------------------------------------------------------------
constexpr size_t kNoAllocRowsBytes = size_t(kNoAllocMaxN) * (kNoAllocMaxN / 2 + 1) *
                                     sizeof(cplx<double>);
constexpr size_t kNoAllocTw2Offset = kNoAllocRowsBytes;
constexpr size_t kNoAllocTw3Offset = kNoAllocTw2Offset + size_t(kTwPow2) * sizeof(cplx<double>);
constexpr size_t kNoAllocBytes = kNoAllocTw3Offset + size_t(kTwPow3) * sizeof(cplx<double>);

namespace noalloc {
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
}

namespace {
// exp(+2 pi i m / n), m = 0..n-1, rounded from long double (the values the per-size twiddle
// cache holds: every table here has a power-of-two ratio to the sizes it serves).
std::vector<cplx<double>> host_twiddles(int n)
{
    std::vector<cplx<double>> h(n);
    for (int k = 0; k < n; ++k) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * k / n;
        h[k].x = double(std::cos(ang));
        h[k].y = double(std::sin(ang));
    }
    return h;
}
}  // namespace

// Device address of this device's copy of noalloc::scratch.  The first call on a device looks
// the address up and writes the twiddle tables into it (synchronous copies from host memory,
// no allocation); later calls return the cached address.
int noalloc_base(unsigned char** base)
{
    static std::mutex m;
    static unsigned char* ptrs[64] = {};
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) {
        set_error("gsmi_irfft_f64_noalloc: device index out of range");
        return GSMI_E_ARG;
    }
    std::lock_guard<std::mutex> lock(m);
    if (!ptrs[dev]) {
        void* p = nullptr;
        rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc::scratch)),
                       "hipGetSymbolAddress(static scratch)");
        if (rc) return rc;
        unsigned char* b = static_cast<unsigned char*>(p);
        const std::vector<cplx<double>> t2 = host_twiddles(kTwPow2);
        const std::vector<cplx<double>> t3 = host_twiddles(kTwPow3);
        rc = hip_check(hipMemcpy(b + kNoAllocTw2Offset, t2.data(), t2.size() * sizeof(t2[0]),
                                 hipMemcpyHostToDevice), "static twiddle upload");
        if (rc) return rc;
        rc = hip_check(hipMemcpy(b + kNoAllocTw3Offset, t3.data(), t3.size() * sizeof(t3[0]),
                                 hipMemcpyHostToDevice), "static twiddle upload");
        if (rc) return rc;
        ptrs[dev] = b;
    }
    *base = ptrs[dev];
    return GSMI_OK;
}

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
    Tables<float> tb{tables, tables32, reinterpret_cast<const float4*>(tables32)};
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

int gsmi_apply_kimage_phases_f64(double* kimage, int xmin, int ymin, int nx, int ny, double dk,
                                 const double* jac, double cenx, double ceny, double flux_scaling,
                                 void* stream)
{
    if (!kimage || nx <= 0 || ny <= 0) {
        set_error("gsmi_apply_kimage_phases_f64: bad arguments");
        return GSMI_E_ARG;
    }
    const double A = jac ? jac[0] : 1., B = jac ? jac[1] : 0., C = jac ? jac[2] : 0.,
                 D = jac ? jac[3] : 1.;
    const int total = nx * ny;
    hipLaunchKernelGGL(apply_kimage_phases_f64, dim3((total + kThreads - 1) / kThreads),
                       dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<cplx<double>*>(kimage), xmin, ymin, nx, ny, dk, A, B, C,
                       D, cenx, ceny, flux_scaling);
    return hip_check(hipGetLastError(), "apply_kimage_phases launch");
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
    const cplx<double>* tw = twiddles<double>(N);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return GSMI_E_HIP;
    }
    return irfft_launch(kin, N, out, reinterpret_cast<cplx<double>*>(scratch), tw, 1, stream);
}

int gsmi_irfft_f64_noalloc(const double* kin, int N, double* out, void* stream)
{
    if (!kin || !out) {
        set_error("gsmi_irfft_f64_noalloc: null buffer");
        return GSMI_E_ARG;
    }
    if (N > kNoAllocMaxN || !fft_size_ok(N, sizeof(cplx<double>))) {
        set_error("gsmi_irfft_f64_noalloc: unsupported FFT size " + std::to_string(N) +
                  " (need 2^k or 3*2^k, <= " + std::to_string(kNoAllocMaxN) +
                  ", the size of its static scratch)");
        return GSMI_E_SIZE;
    }
    unsigned char* base = nullptr;
    int rc = noalloc_base(&base);
    if (rc) return rc;
    const bool p3 = (N % 3 == 0);
    const cplx<double>* tw = reinterpret_cast<const cplx<double>*>(
        base + (p3 ? kNoAllocTw3Offset : kNoAllocTw2Offset));
    const int tws = p3 ? kTwPow3 / N : kTwPow2 / N;
    return irfft_launch(kin, N, out, reinterpret_cast<cplx<double>*>(base), tw, tws, stream);
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
    Tables<float> tb{tables, tables32, reinterpret_cast<const float4*>(tables32)};
    return fold_impl<float>(stamps, nstamps, N, ops, leaves, tb, scratch, stream);
}

int gsmi_row_c2r_f64(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, const void* scratch,
                     double* out, void* stream)
{
    if (!stamps || !scratch || !out || max_ny <= 0) {
        set_error("bad argument");
        return GSMI_E_ARG;
    }
    return rows_impl<double>(stamps, nstamps, N, max_ny, scratch, out, stream);
}

int gsmi_row_c2r_f32(const gsmi_stamp* stamps, int nstamps, int N, int max_ny, const void* scratch,
                     float* out, void* stream)
{
    if (!stamps || !scratch || !out || max_ny <= 0) {
        set_error("bad argument");
        return GSMI_E_ARG;
    }
    return rows_impl<float>(stamps, nstamps, N, max_ny, scratch, out, stream);
}

int gsmi_fold_list_f64(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const int32_t* cliprows, void* scratch, void* stream)
{
    if (!ops || !leaves || !scratch) { set_error("null buffer"); return GSMI_E_ARG; }
    Tables<double> tb{tables, cliprows};
    return fold_list_impl<double>(stamps, nstamps, sched, ops, leaves, tb, scratch, stream);
}

int gsmi_fold_list_f32(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const float* tables32, void* scratch, void* stream)
{
    if (!ops || !leaves || !scratch) { set_error("null buffer"); return GSMI_E_ARG; }
    Tables<float> tb{tables, tables32, reinterpret_cast<const float4*>(tables32)};
    return fold_list_impl<float>(stamps, nstamps, sched, ops, leaves, tb, scratch, stream);
}

int gsmi_rows_list_f64(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const void* scratch, double* out, void* stream)
{
    if (!scratch || !out) { set_error("null buffer"); return GSMI_E_ARG; }
    return rows_list_impl<double>(stamps, nstamps, sched, scratch, out, stream);
}

int gsmi_rows_list_f32(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const void* scratch, float* out, void* stream)
{
    if (!scratch || !out) { set_error("null buffer"); return GSMI_E_ARG; }
    return rows_list_impl<float>(stamps, nstamps, sched, scratch, out, stream);
}

int gsmi_draw_list_f64(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const int32_t* cliprows, void* scratch, double* out, void* stream)
{
    int rc = gsmi_fold_list_f64(stamps, nstamps, sched, ops, leaves, tables, cliprows, scratch,
                                stream);
    if (rc) return rc;
    return gsmi_rows_list_f64(stamps, nstamps, sched, scratch, out, stream);
}

int gsmi_draw_list_f32(const gsmi_stamp* stamps, int nstamps, const gsmi_schedule* sched,
                       const gsmi_op* ops, const gsmi_leaf* leaves, const double* tables,
                       const float* tables32, void* scratch, float* out, void* stream)
{
    int rc = gsmi_fold_list_f32(stamps, nstamps, sched, ops, leaves, tables, tables32, scratch,
                                stream);
    if (rc) return rc;
    return gsmi_rows_list_f32(stamps, nstamps, sched, scratch, out, stream);
}

}
