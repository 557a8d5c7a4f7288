// qed2d_kernels.hip -- HIP kernels of the QED2d molecular-dynamics path (gfx950).
//
// Every kernel here is HBM-bandwidth bound (the Wilson stencil does ~0.6
// flop/byte in FP64), so the design goals are: one pass over each field per
// logical use, coalesced 16-byte (FP64) or 8-byte (FP32) accesses along i1,
// vertical neighbours kept in registers by marching each thread up a column
// of `R` rows, and reductions finished on the device by the last block to
// arrive so that no kernel boundary needs a host round trip.
#include <type_traits>

#include "qed2d_fields.h"

namespace qed2d {

// ===========================================================================
// Launch geometry of the marching stencil kernels.
// A block holds kBlock threads arranged bx (along i1) x by (along i2).  Each
// thread owns one i1 column and marches up R consecutive rows.
// ===========================================================================
MarchGrid march_grid(const Latt &L, int rows)
{
    MarchGrid g;
    int bx = ((L.L1 + kWave - 1) / kWave) * kWave;
    if (bx > kBlock) bx = kBlock;
    g.bx = bx;
    g.by = kBlock / bx;
    g.rows = rows;
    g.nbx = (L.L1 + bx - 1) / bx;
    int rowgroups = (L.L2 + rows - 1) / rows;
    g.nby = (rowgroups + g.by - 1) / g.by;
    return g;
}

// ===========================================================================
// Q apply: so = gamma5 Dw si  (src/QED2dDirac.jl:2-33)
// ===========================================================================
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_q_apply(Latt L, const c2<T> *__restrict__ U,
                                                    const c2<T> *__restrict__ si,
                                                    c2<T> *__restrict__ so, T mass2, int bx,
                                                    int by, int nbx)
{
    MarchPos p = march_pos(L, bx, by, nbx, R, apply_bands(L));
    if (!p.active) return;
    SrcPlain<T> src{si, L.V};
    EpiStore<T> epi{so, L.V};
    q_march<T, R>(L, U, mass2, src, epi, p.i1, p.r0);
}

// so = Q si + mu^2 x
struct EpiShift {
    double2 *__restrict__ so;
    const double2 *__restrict__ x;
    double mu2;
    long V;
    struct Pre {
        double2 x1, x2;
    };
    __device__ __forceinline__ Pre pre(long s) const { return Pre{x[s], x[V + s]}; }
    __device__ __forceinline__ void operator()(long s, const Pre &q, double2 o1, double2 o2,
                                               double2, double2)
    {
        so[s] = mk<double>(o1.x + mu2 * q.x1.x, o1.y + mu2 * q.x1.y);
        so[V + s] = mk<double>(o2.x + mu2 * q.x2.x, o2.y + mu2 * q.x2.y);
    }
};

template <int R>
__global__ void __launch_bounds__(kBlock) k_q_shift(Latt L, const double2 *__restrict__ U,
                                                    const double2 *__restrict__ si,
                                                    const double2 *__restrict__ x,
                                                    double2 *__restrict__ so, double mass2,
                                                    double mu2, int bx, int by, int nbx)
{
    MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    SrcPlain<double> src{si, L.V};
    EpiShift epi{so, x, mu2, L.V};
    q_march<double, R>(L, U, mass2, src, epi, p.i1, p.r0);
}


// FP32 apply, two sites per thread along i1.
template <int R>
__global__ void __launch_bounds__(kBlock) k_q_apply2(Latt L, const float2 *__restrict__ U,
                                                     const float2 *__restrict__ si,
                                                     float2 *__restrict__ so, float mass2, int bx,
                                                     int by, int nbx)
{
    MarchPos p = march_pos2(L, bx, by, nbx, R, apply_bands(L));
    if (!p.active) return;
    SrcPlain2 src{si, L.V};
    EpiStore2 epi{so, L.V};
    q_march2<R>(L, U, mass2, src, epi, p.i1, p.r0);
}

template <typename T>
static void q_apply_impl(const Latt &L, const c2<T> *U, const c2<T> *si, c2<T> *so, T am0,
                         int rows, hipStream_t st)
{
    if constexpr (std::is_same<T, float>::value) {
        if (pair_ok(L)) {
            MarchGrid g = march_grid2(L, rows);
            dim3 grid(g.nblocks()), blk(kBlock);
            DISPATCH_ROWS(rows, k_q_apply2, grid, blk, 0, st, L, U, si, so, T(2) + am0, g.bx, g.by,
                          g.nbx);
            return;
        }
    }
    MarchGrid g = march_grid(L, rows);
    const T mass2 = T(2) + am0;
    dim3 grid(g.nblocks()), blk(kBlock);
    switch (rows) {
    case 1: launch_k((k_q_apply<T, 1>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    case 2: launch_k((k_q_apply<T, 2>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    case 4: launch_k((k_q_apply<T, 4>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    case 8: launch_k((k_q_apply<T, 8>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    default: launch_k((k_q_apply<T, 16>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    }
}

void launch_q_apply_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                        double am0, int rows, hipStream_t st)
{
    q_apply_impl<double>(L, U, si, so, am0, rows, st);
}
void launch_q_apply_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so, float am0,
                        int rows, hipStream_t st)
{
    q_apply_impl<float>(L, U, si, so, am0, rows, st);
}

void launch_q_shift_f64(const Latt &L, const double2 *U, const double2 *si, const double2 *x,
                        double2 *so, double am0, double mu2, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    DISPATCH_ROWS(rows, k_q_shift, grid, blk, 0, st, L, U, si, x, so, 2.0 + am0, mu2, g.bx, g.by,
                  g.nbx);
}

// ---------------------------------------------------------------------------
// Q apply to k right-hand sides sharing the links: the links of a row are
// loaded once and reused for every RHS (32 + 64k bytes per site in FP64).
// Layout: RHS j occupies si[j * 2V .. (j+1) * 2V).
// ---------------------------------------------------------------------------
// k right-hand sides sharing the links.  A tile is 64 columns (one wave) by
// R rows; the 4 waves of a block march the same tile for 4 consecutive RHS
// (their link loads coincide in the CU's L1), and the logical block index t
// (grouped by XCD) maps to RHS group t % ngrp, tile t / ngrp, so every RHS of
// a tile runs on one XCD and its links come from HBM once.  (One RHS per block
// over a 256-column tile: 222 / 116 us FP64 / FP32 at 1024^2 x 16; this form
// 217 / 113 us.)
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_q_mrhs(Latt L, const c2<T> *__restrict__ U,
                                                   const c2<T> *__restrict__ si,
                                                   c2<T> *__restrict__ so, int nrhs, T mass2,
                                                   int nbx, int ngrp)
{
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int g = t % ngrp, tile = t / ngrp;
    const int j = g * (kBlock / kWave) + threadIdx.x / kWave;
    const int i1 = (tile % nbx) * kWave + threadIdx.x % kWave;
    const int r0 = (tile / nbx) * R;
    if (j >= nrhs || i1 >= L.L1 || r0 >= L.L2) return;
    SrcPlain<T> src{si + (long)j * 2 * L.V, L.V};
    EpiStore<T> epi{so + (long)j * 2 * L.V, L.V};
    q_march<T, R>(L, U, mass2, src, epi, i1, r0);
}

template <int R>
__global__ void __launch_bounds__(kBlock) k_q_mrhs2(Latt L, const float2 *__restrict__ U,
                                                    const float2 *__restrict__ si,
                                                    float2 *__restrict__ so, int nrhs,
                                                    float mass2, int nbx, int ngrp)
{
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int g = t % ngrp, tile = t / ngrp;
    const int j = g * (kBlock / kWave) + threadIdx.x / kWave;
    const int jp = (tile % nbx) * kWave + threadIdx.x % kWave;
    const int r0 = (tile / nbx) * R;
    if (j >= nrhs || jp >= L.L1 / 2 || r0 >= L.L2) return;
    SrcPlain2 src{si + (long)j * 2 * L.V, L.V};
    EpiStore2 epi{so + (long)j * 2 * L.V, L.V};
    q_march2<R>(L, U, mass2, src, epi, jp, r0);
}

template <typename T>
static void q_mrhs_impl(const Latt &L, const c2<T> *U, const c2<T> *si, c2<T> *so, int nrhs,
                        T am0, int rows, hipStream_t st)
{
    const T m2 = T(2) + am0;
    const int ngrp = (nrhs + kBlock / kWave - 1) / (kBlock / kWave);
    const int nby = (L.L2 + rows - 1) / rows;
    if constexpr (std::is_same<T, float>::value) {
        if (pair_ok(L)) {
            const int nbx = (L.L1 / 2 + kWave - 1) / kWave;
            dim3 grid(nbx * nby * ngrp), blk(kBlock);
            DISPATCH_ROWS(rows, k_q_mrhs2, grid, blk, 0, st, L, U, si, so, nrhs, m2, nbx, ngrp);
            return;
        }
    }
    const int nbx = (L.L1 + kWave - 1) / kWave;
    dim3 grid(nbx * nby * ngrp), blk(kBlock);
    switch (rows) {
    case 1: launch_k((k_q_mrhs<T, 1>), grid, blk, 0, st, L, U, si, so, nrhs, m2, nbx, ngrp); break;
    case 2: launch_k((k_q_mrhs<T, 2>), grid, blk, 0, st, L, U, si, so, nrhs, m2, nbx, ngrp); break;
    case 4: launch_k((k_q_mrhs<T, 4>), grid, blk, 0, st, L, U, si, so, nrhs, m2, nbx, ngrp); break;
    case 8: launch_k((k_q_mrhs<T, 8>), grid, blk, 0, st, L, U, si, so, nrhs, m2, nbx, ngrp); break;
    default: launch_k((k_q_mrhs<T, 16>), grid, blk, 0, st, L, U, si, so, nrhs, m2, nbx, ngrp); break;
    }
}
void launch_q_apply_mrhs_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                             int nrhs, double am0, int rows, hipStream_t st)
{
    q_mrhs_impl<double>(L, U, si, so, nrhs, am0, rows, st);
}
void launch_q_apply_mrhs_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so,
                             int nrhs, float am0, int rows, hipStream_t st)
{
    q_mrhs_impl<float>(L, U, si, so, nrhs, am0, rows, st);
}

// ===========================================================================
// STREAM kernels (achievable HBM bandwidth of the device, 16-byte elements):
// copy b = a, triad c = a + s b, read sum(a) into per-block partials.
// ===========================================================================
__global__ void __launch_bounds__(kBlock) k_stream_copy(const double2 *__restrict__ a,
                                                        double2 *__restrict__ b, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}
__global__ void __launch_bounds__(kBlock) k_stream_triad(const double2 *__restrict__ a,
                                                         const double2 *__restrict__ b,
                                                         double2 *__restrict__ c, double s,
                                                         long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 x = a[i], y = b[i];
        c[i] = mk<double>(x.x + s * y.x, x.y + s * y.y);
    }
}
__global__ void __launch_bounds__(kBlock) k_stream_read(const double2 *__restrict__ a,
                                                        double *__restrict__ part, long n)
{
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double2 x = a[i];
        acc += x.x + x.y;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}
// k arrays of m elements each (back to back) copied into k others: block t
// copies 256 elements of array t % k, so consecutive blocks stream from k
// places at once.
__global__ void __launch_bounds__(kBlock) k_stream_copy_k(const double2 *__restrict__ a,
                                                          double2 *__restrict__ b, long m, int k)
{
    const long t = blockIdx.x;
    const long e = (t / k) * kBlock + threadIdx.x;
    if (e < m) b[(t % k) * m + e] = a[(t % k) * m + e];
}
void launch_stream(int which, const double2 *a, const double2 *b, double2 *c, double *part,
                   long n, hipStream_t st)
{
    const long nb = (n + kBlock - 1) / kBlock;
    if (which == 3) {
        const long m = n / 16;  // (n: a multiple of 16 x 256)
        launch_k(k_stream_copy_k, dim3(16 * ((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                 a, c, m, 16);
        return;
    }
    if (which == 0)
        launch_k(k_stream_copy, dim3(nb), dim3(kBlock), 0, st, a, c, n);
    else if (which == 1)
        launch_k(k_stream_triad, dim3(nb), dim3(kBlock), 0, st, a, b, c, 0.5, n);
    else
        launch_k(k_stream_read, dim3(kStreamReadBlocks), dim3(kBlock), 0, st, a, part,
                           n);
}

// ===========================================================================
// Elementwise helpers (grid-stride).
// ===========================================================================
static inline int ew_blocks(long n)
{
    long b = (n + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

__global__ void k_gamm5(const double2 *__restrict__ si, double2 *__restrict__ so, long V)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < V;
         i += (long)gridDim.x * blockDim.x) {
        so[i] = si[i];
        so[V + i] = cneg<double>(si[V + i]);
    }
}
void launch_gamm5(const Latt &L, const double2 *si, double2 *so, hipStream_t st)
{
    launch_k(k_gamm5, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, si, so, L.V);
}

__global__ void k_axpy_c(double2 *__restrict__ y, const double2 *__restrict__ x, double ar,
                         double ai, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 a = mk<double>(ar, ai);
        y[i] = cadd<double>(y[i], cmul<double>(a, x[i]));
    }
}
void launch_axpy_c(double2 *y, const double2 *x, double ar, double ai, long n, hipStream_t st)
{
    launch_k(k_axpy_c, dim3(ew_blocks(n)), dim3(kBlock), 0, st, y, x, ar, ai, n);
}

// y = a y + b x (x may alias y)
__global__ void k_scale_add(double2 *y, double a, const double2 *x, double b, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 yv = y[i], xv = x[i];
        y[i] = mk<double>(a * yv.x + b * xv.x, a * yv.y + b * xv.y);
    }
}
void launch_scale_add(double2 *y, double a, const double2 *x, double b, long n, hipStream_t st)
{
    launch_k(k_scale_add, dim3(ew_blocks(n)), dim3(kBlock), 0, st, y, a, x, b, n);
}

// y = a y + b x with one coefficient (which = 0: a, 1: b) replaced by
// 1 / red->v[0], the result of a reduction still on the device: the power
// method's normalisation without a host read.  The same expression as
// k_scale_add (and 1 / v[0] rounds as on the host).
__global__ void k_scale_add_inv(double2 *y, double a, const double2 *x, double b,
                                const RedOut *red, int which, long n)
{
    const double s = 1.0 / red->v[0];
    if (which == 0) a = s;
    else b = s;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 yv = y[i], xv = x[i];
        y[i] = mk<double>(a * yv.x + b * xv.x, a * yv.y + b * xv.y);
    }
}
void launch_scale_add_inv(double2 *y, double a, const double2 *x, double b, const RedOut *red,
                          int which, long n, hipStream_t st)
{
    launch_k(k_scale_add_inv, dim3(ew_blocks(n)), dim3(kBlock), 0, st, y, a, x, b, red, which,
             n);
}

__global__ void k_copy(double2 *__restrict__ d, const double2 *__restrict__ s, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        d[i] = s[i];
}
void launch_copy(double2 *dst, const double2 *src, long n, hipStream_t st)
{
    launch_k(k_copy, dim3(ew_blocks(n)), dim3(kBlock), 0, st, dst, src, n);
}

__global__ void k_negate(double *a, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        a[i] = -a[i];
}
void launch_negate(double *a, long n, hipStream_t st)
{
    launch_k(k_negate, dim3(ew_blocks(n)), dim3(kBlock), 0, st, a, n);
}


// ===========================================================================
// Deterministic reductions without in-kernel grid synchronisation: a
// producer kernel writes one partial per block, a one-block finishing kernel
// sums them in block order.  (A grid-wide "last block" finish needs an
// agent-scope release fence per block, which on a multi-XCD part writes back
// every dirty L2 line.)
// ===========================================================================
constexpr int kRedBlocks = 1024;

static inline int red_blocks(long n)
{
    long b = (n + kBlock - 1) / kBlock;
    if (b > kRedBlocks) b = kRedBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

// a^dag b (complex) or |a|^2 when b == nullptr: partials (re, im) per block.
__global__ void __launch_bounds__(kBlock) k_dot(const double2 *__restrict__ a,
                                                const double2 *__restrict__ b, long n,
                                                double *part)
{
    double sr = 0.0, si = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 x = a[i];
        if (b) {
            double2 y = b[i];
            sr += x.x * y.x + x.y * y.y;
            si += x.x * y.y - x.y * y.x;
        } else {
            sr += x.x * x.x + x.y * x.y;
        }
    }
    double br, bi;
    block_sum2_t0(sr, si, br, bi);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = br;
        part[2 * blockIdx.x + 1] = bi;
    }
}

// Sums nv interleaved plain partials: out->v[k] = sum_b part[b * nv + k].
__global__ void __launch_bounds__(kBlock) k_finish_sum(const double *part, int nblocks, int nv,
                                                       RedOut *out)
{
    __shared__ double lds[kBlock / kWave + 1];
    for (int k = 0; k < nv; ++k) {
        double s = sum_prev_partials(part, nblocks, nv, k, lds);
        if (threadIdx.x == 0) out->v[k] = s;
    }
}

void launch_norm2(const double2 *a, long n, RedOut *out, double *part, hipStream_t st)
{
    const int nb = red_blocks(n);
    launch_k(k_dot, dim3(nb), dim3(kBlock), 0, st, a, nullptr, n, part);
    launch_k(k_finish_sum, dim3(1), dim3(kBlock), 0, st, part, nb, 2, out);
}
void launch_dot(const double2 *a, const double2 *b, long n, RedOut *out, double *part,
                hipStream_t st)
{
    const int nb = red_blocks(n);
    launch_k(k_dot, dim3(nb), dim3(kBlock), 0, st, a, b, n, part);
    launch_k(k_finish_sum, dim3(1), dim3(kBlock), 0, st, part, nb, 2, out);
}

// Block-wide double-double sum; result in thread 0.
__device__ __forceinline__ dd block_sum_dd(dd v, double *lds)
{
    v = wave_sum_dd(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) {
        lds[2 * w] = v.hi;
        lds[2 * w + 1] = v.lo;
    }
    __syncthreads();
    dd s = {0.0, 0.0};
    if (threadIdx.x == 0) {
        for (int i = 0; i < nw; ++i) {
            dd t = {lds[2 * i], lds[2 * i + 1]};
            s = dd_add(s, t);
        }
    }
    __syncthreads();
    return s;
}

// Sum Re P, sum arg P (src/QED2dAction.jl:2-32) and sum p^2 (src/QED2dHMC.jl:200),
// each as a double-double, for block blk of nblk over the lattice (grid-stride
// over its sites); optional per-site density outputs.  Thread 0 writes the
// block's partials to pp[0..5].
__device__ __forceinline__ void observables_blk(const Latt &L, const double2 *__restrict__ U,
                                                const double *__restrict__ mom,
                                                double *__restrict__ plx,
                                                double *__restrict__ qdens, double *pp, int blk,
                                                int nblk, double *lds)
{
    dd sp = {0.0, 0.0}, sq = {0.0, 0.0}, sm = {0.0, 0.0};
    for (long s = blk * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)nblk * blockDim.x) {
        const int i2 = (int)(s / L.L1), i1 = (int)(s - (long)i2 * L.L1);
        double2 P = plaq_at(U, L, i1, i2);
        double q = atan2(P.y, P.x);
        sp = dd_add(sp, P.x);
        sq = dd_add(sq, q);
        if (plx) plx[s] = P.x;
        if (qdens) qdens[s] = q;
        if (mom) {
            double m0 = mom[s], m1 = mom[L.V + s];
            sm = dd_add(sm, m0 * m0);
            sm = dd_add(sm, m1 * m1);
        }
    }
    dd bp = block_sum_dd(sp, lds);
    dd bq = block_sum_dd(sq, lds);
    dd bm = block_sum_dd(sm, lds);
    if (threadIdx.x == 0) {
        pp[0] = bp.hi; pp[1] = bp.lo; pp[2] = bq.hi; pp[3] = bq.lo; pp[4] = bm.hi; pp[5] = bm.lo;
    }
}
__global__ void __launch_bounds__(kBlock) k_observables(Latt L, const double2 *__restrict__ U,
                                                        const double *__restrict__ mom,
                                                        double *__restrict__ plx,
                                                        double *__restrict__ qdens,
                                                        double *part)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    observables_blk(L, U, mom, plx, qdens, part + 6 * (long)blockIdx.x, blockIdx.x, gridDim.x,
                    lds);
}

// Double-double finish of nq (hi, lo) partial pairs of nblocks blocks:
// out->v[2k], v[2k+1].
__device__ __forceinline__ void finish_dd_blk(const double *part, int nblocks, int nq,
                                              RedOut *out, double *lds)
{
    for (int k = 0; k < nq; ++k) {
        dd acc = {0.0, 0.0};
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
            dd t = {part[(long)b * 2 * nq + 2 * k], part[(long)b * 2 * nq + 2 * k + 1]};
            acc = dd_add(acc, t);
        }
        acc = block_sum_dd(acc, lds);
        if (threadIdx.x == 0) {
            out->v[2 * k] = acc.hi;
            out->v[2 * k + 1] = acc.lo;
        }
    }
}
__global__ void __launch_bounds__(kBlock) k_finish_dd(const double *part, int nblocks, int nq,
                                                      RedOut *out)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    finish_dd_blk(part, nblocks, nq, out, lds);
}

// a^dag b or |a|^2 (b == nullptr) over n elements accumulated as double-doubles,
// block blk of nblk: partials (re hi, re lo, im hi, im lo) to pp[0..3].  For the
// fermion terms of H, whose size is ~V and whose differences between two
// trajectories must be resolved below 1e-9.
__device__ __forceinline__ void dot_dd_blk(const double2 *__restrict__ a,
                                           const double2 *__restrict__ b, long n, double *pp,
                                           int blk, int nblk, double *lds)
{
    dd sr = {0.0, 0.0}, si = {0.0, 0.0};
    for (long i = blk * (long)blockDim.x + threadIdx.x; i < n; i += (long)nblk * blockDim.x) {
        const double2 x = a[i];
        if (b) {
            const double2 y = b[i];
            sr = dd_add(sr, x.x * y.x);
            sr = dd_add(sr, x.y * y.y);
            si = dd_add(si, x.x * y.y);
            si = dd_add(si, -(x.y * y.x));
        } else {
            sr = dd_add(sr, x.x * x.x);
            sr = dd_add(sr, x.y * x.y);
        }
    }
    const dd br = block_sum_dd(sr, lds);
    const dd bi = block_sum_dd(si, lds);
    if (threadIdx.x == 0) {
        pp[0] = br.hi; pp[1] = br.lo; pp[2] = bi.hi; pp[3] = bi.lo;
    }
}
__global__ void __launch_bounds__(kBlock) k_dot_dd(const double2 *__restrict__ a,
                                                   const double2 *__restrict__ b, long n,
                                                   double *part)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    dot_dd_blk(a, b, n, part + 4 * (long)blockIdx.x, blockIdx.x, gridDim.x, lds);
}

// ---------------------------------------------------------------------------
// The same reductions over the chains of a chain set, one launch for all:
// chain c owns blocks blk0[c] .. blk0[c+1]-1, as many as its single-lattice
// reduction launches (red_blocks), each doing what that launch's block of the
// same index does, so every chain's sums carry the single-lattice bits.  The
// finishing kernel runs one block per chain.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int chain_of_block(const int *blk0, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__global__ void __launch_bounds__(kBlock) kb_observables(const BChain *ch, const int *blk0, int n,
                                                         const double2 *__restrict__ U,
                                                         const double *__restrict__ mom,
                                                         double *part)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    const int c = chain_of_block(blk0, n, blockIdx.x);
    const BChain &b = ch[c];
    observables_blk(b.L, U + 2 * b.uoff, mom ? mom + 2 * b.off : nullptr, nullptr, nullptr,
                    part + 6 * (long)blockIdx.x, blockIdx.x - blk0[c], blk0[c + 1] - blk0[c],
                    lds);
}
__global__ void __launch_bounds__(kBlock) kb_dot_dd(const BChain *ch, const int *blk0, int n,
                                                    const double2 *__restrict__ a,
                                                    const double2 *__restrict__ b,
                                                    double *part)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    const int c = chain_of_block(blk0, n, blockIdx.x);
    const long o = 2 * ch[c].off;
    dot_dd_blk(a + o, b ? b + o : nullptr, 2 * ch[c].L.V, part + 4 * (long)blockIdx.x,
               blockIdx.x - blk0[c], blk0[c + 1] - blk0[c], lds);
}
__global__ void __launch_bounds__(kBlock) kb_finish_dd(const int *blk0, const double *part,
                                                       int nq, RedOut *out)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    const int c = blockIdx.x;
    finish_dd_blk(part + 2L * nq * blk0[c], blk0[c + 1] - blk0[c], nq, out + c, lds);
}
int chain_red_blocks(long n) { return red_blocks(n); }
void launch_b_observables(const BChain *ch, const int *blk0, int n, int nblocks,
                          const double2 *U, const double *mom, double *part, RedOut *out,
                          hipStream_t st)
{
    launch_k(kb_observables, dim3(nblocks), dim3(kBlock), 0, st, ch, blk0, n, U, mom, part);
    launch_k(kb_finish_dd, dim3(n), dim3(kBlock), 0, st, blk0, (const double *)part, 3, out);
}
void launch_b_dot_dd(const BChain *ch, const int *blk0, int n, int nblocks, const double2 *a,
                     const double2 *b, double *part, RedOut *out, hipStream_t st)
{
    launch_k(kb_dot_dd, dim3(nblocks), dim3(kBlock), 0, st, ch, blk0, n, a, b, part);
    launch_k(kb_finish_dd, dim3(n), dim3(kBlock), 0, st, blk0, (const double *)part, 2, out);
}

void launch_dot_dd(const double2 *a, const double2 *b, long n, RedOut *out, double *part,
                   hipStream_t st)
{
    const int nb = red_blocks(n);
    launch_k(k_dot_dd, dim3(nb), dim3(kBlock), 0, st, a, b, n, part);
    launch_k(k_finish_dd, dim3(1), dim3(kBlock), 0, st, part, nb, 2, out);
}

void launch_observables(const Latt &L, const double2 *U, const double *mom, double *plx,
                        double *qdens, RedOut *out, double *part, hipStream_t st)
{
    const int nb = red_blocks(L.V);
    launch_k(k_observables, dim3(nb), dim3(kBlock), 0, st, L, U, mom, plx, qdens, part);
    launch_k(k_finish_dd, dim3(1), dim3(kBlock), 0, st, part, nb, 3, out);
}
// ===========================================================================
// Forces and link updates
// ===========================================================================

__global__ void k_tr_dQwdU(Latt L, const double2 *__restrict__ U, const double2 *__restrict__ X,
                           const double2 *__restrict__ Y, double *__restrict__ frc)
{
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)gridDim.x * blockDim.x) {
        const int i2 = (int)(s / L.L1), i1 = (int)(s - (long)i2 * L.L1);
        double f1, f2;
        ferm_force(U, X, Y, L, i1, i2, f1, f2);
        frc[s] = f1;
        frc[L.V + s] = f2;
    }
}
void launch_tr_dQwdU(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                     double *frc, hipStream_t st)
{
    launch_k(k_tr_dQwdU, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, U, X, Y, frc);
}

// Scatter form, as the reference kernel (src/QED2dAction.jl:34-53).
__global__ void k_krnl_force(Latt L, const double2 *__restrict__ U, double beta,
                             double *__restrict__ frc1, double *__restrict__ frc2)
{
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)gridDim.x * blockDim.x) {
        const int i2 = (int)(s / L.L1), i1 = (int)(s - (long)i2 * L.L1);
        const int iu1 = (i1 + 1 == L.L1) ? 0 : i1 + 1;
        const int iu2 = (i2 + 1 == L.L2) ? 0 : i2 + 1;
        const double v = beta * plaq_at(U, L, i1, i2).y;
        frc1[s] = -v;
        frc1[L.V + s] = v;
        frc2[L.V + (long)i2 * L.L1 + iu1] = -v;
        frc2[(long)iu2 * L.L1 + i1] = v;
    }
}
void launch_krnl_force(const Latt &L, const double2 *U, double beta, double *frc1, double *frc2,
                       hipStream_t st)
{
    launch_k(k_krnl_force, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, U, beta, frc1,
                       frc2);
}

// P = U1(x) U2(x+1) conj(U1(x+2) U2(x)) from the four links, the expression
// of plaq_at.
__device__ __forceinline__ double2 plaq4(double2 u1, double2 u2r, double2 u1n, double2 u2)
{
    const double2 a = cmul<double>(u1, u2r);
    const double2 b = cmul<double>(u1n, u2);
    return cmul<double>(a, cconj<double>(b));
}

// The gauge force alone (the gather form of src/QED2dAction.jl:34-53) and the
// momentum update, mom += eps (F_gauge + facc), marching rows as the stencil
// does: a thread owns column i1 and rows r0 .. r0+R-1, carries v = beta Im P
// of the row below (P(x-2)), and takes U2(x+1) and v(x-1) from the adjacent
// lanes, loading or computing them at the wave's edges and the periodic wrap.
// Each link is loaded once, plus one halo row per tile; per site
// g1 = -v(x) + v(x-2), g2 = v(x) - v(x-1) as force_mom_site.
template <int R>
__global__ void __launch_bounds__(kBlock) k_gauge_mom(Latt L, const double2 *__restrict__ U,
                                                      const double *__restrict__ facc,
                                                      double *__restrict__ mom, double beta,
                                                      double eps, int bx, int by, int nbx)
{
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int i1 = p.i1, r0 = p.r0;
    const int i1p = (i1 + 1 == L1) ? 0 : i1 + 1;
    const int i1m = (i1 == 0) ? L1 - 1 : i1 - 1;
    const int lane = threadIdx.x & (kWave - 1);
    const bool ldR = (lane == kWave - 1) || (i1p == 0);  // U2(x+1) not in lane+1
    const bool ldL = (lane == 0) || (i1 == 0);           // v(x-1) not in lane-1
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    const int nr = (r0 + R <= L2) ? R : L2 - r0;
    const long sm = (long)rm * L1, s0 = (long)r0 * L1;
    // the row below the tile: its plaquette is P(x-2) of row r0
    double2 u1 = U[sm + i1], u2 = U[V + sm + i1];
    double2 u1n = U[s0 + i1];
    double2 u2r = from_right<double>(u2);
    if (ldR) u2r = U[V + sm + i1p];
    double vprev = beta * plaq4(u1, u2r, u1n, u2).y;
    double2 l1{};  // U1(x-1) of the current row (edge lanes)
    if (ldL) l1 = U[s0 + i1m];
    u1 = u1n;
    u2 = U[V + s0 + i1];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k >= nr) break;
        const int r = r0 + k;
        const int rn = (r + 1 == L2) ? 0 : r + 1;
        const long row = (long)r * L1, s = row + i1, sn = (long)rn * L1;
        u1n = U[sn + i1];
        double2 u2n{};
        if (k + 1 < nr) u2n = U[V + sn + i1];
        double m1 = mom[s], m2 = mom[V + s];
        u2r = from_right<double>(u2);
        if (ldR) u2r = U[V + row + i1p];
        const double v = beta * plaq4(u1, u2r, u1n, u2).y;
        double vl = lane_shift(v, true);
        if (ldL) {
            const double2 l2 = U[V + row + i1m];
            const double2 l1n = U[sn + i1m];
            vl = beta * plaq4(l1, u2, l1n, l2).y;  // P(x-1)
            l1 = l1n;
        }
        double g1 = -v + vprev;
        double g2 = v - vl;
        if (facc) {
            g1 = g1 + facc[s];
            g2 = g2 + facc[V + s];
        }
        mom[s] = m1 + eps * g1;
        mom[V + s] = m2 + eps * g2;
        vprev = v;
        u1 = u1n;
        u2 = u2n;
    }
}

// The Nf=2 force evaluation's gauge + fermion force and momentum update
// (force_mom_site), marching rows as k_gauge_mom, with the fermion fields of the
// thread's column in registers too: X and Y of row r + 1 (x + 2) are loaded
// one row ahead and become the next row's own, those of x + 1 come from the
// adjacent lane (loaded at the wave's right edge and the periodic wrap).  Each
// link and each X, Y element is loaded once, plus one halo row per tile: the
// algorithmic 128 B/site.  Same expressions, same bits as force_mom_site.
template <int R>
__global__ void __launch_bounds__(kBlock) k_force_mom(Latt L, const double2 *__restrict__ U,
                                                        const double2 *__restrict__ X,
                                                        const double2 *__restrict__ Y,
                                                        const double *__restrict__ facc,
                                                        double *__restrict__ mom, double beta,
                                                        double eps, double cf, int bx, int by,
                                                        int nbx)
{
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int i1 = p.i1, r0 = p.r0;
    const int i1p = (i1 + 1 == L1) ? 0 : i1 + 1;
    const int i1m = (i1 == 0) ? L1 - 1 : i1 - 1;
    const int lane = threadIdx.x & (kWave - 1);
    const bool ldR = (lane == kWave - 1) || (i1p == 0);  // x + 1 not in lane+1
    const bool ldL = (lane == 0) || (i1 == 0);           // v(x-1) not in lane-1
    const bool apbc = (L.bc & 1) != 0;
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    const int nr = (r0 + R <= L2) ? R : L2 - r0;
    const long sm = (long)rm * L1, s0 = (long)r0 * L1;
    // the row below the tile: its plaquette is P(x-2) of row r0
    double2 u1 = U[sm + i1], u2 = U[V + sm + i1];
    double2 u1n = U[s0 + i1];
    double2 u2r = from_right<double>(u2);
    if (ldR) u2r = U[V + sm + i1p];
    double vprev = beta * plaq4(u1, u2r, u1n, u2).y;
    double2 l1{};  // U1(x-1) of the current row (edge lanes)
    if (ldL) l1 = U[s0 + i1m];
    u1 = u1n;
    u2 = U[V + s0 + i1];
    double2 x1 = X[s0 + i1], x2 = X[V + s0 + i1], y1 = Y[s0 + i1], y2 = Y[V + s0 + i1];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k >= nr) break;
        const int r = r0 + k;
        const int rn = (r + 1 == L2) ? 0 : r + 1;
        const long row = (long)r * L1, s = row + i1, sn = (long)rn * L1;
        u1n = U[sn + i1];
        double2 u2n{};
        if (k + 1 < nr) u2n = U[V + sn + i1];
        const double2 xb1 = X[sn + i1], xb2 = X[V + sn + i1];
        const double2 yb1 = Y[sn + i1], yb2 = Y[V + sn + i1];
        const double m1 = mom[s], m2 = mom[V + s];
        u2r = from_right<double>(u2);
        if (ldR) u2r = U[V + row + i1p];
        double2 xa1 = from_right<double>(x1), xa2 = from_right<double>(x2);
        double2 ya1 = from_right<double>(y1), ya2 = from_right<double>(y2);
        if (ldR) {
            xa1 = X[row + i1p];
            xa2 = X[V + row + i1p];
            ya1 = Y[row + i1p];
            ya2 = Y[V + row + i1p];
        }
        const double v = beta * plaq4(u1, u2r, u1n, u2).y;
        double vl = lane_shift(v, true);
        if (ldL) {
            const double2 l2 = U[V + row + i1m];
            const double2 l1n = U[sn + i1m];
            vl = beta * plaq4(l1, u2, l1n, l2).y;  // P(x-1)
            l1 = l1n;
        }
        double f1, f2;
        ferm_force_v(u1, (apbc && r == L2 - 1) ? cneg<double>(u2) : u2, x1, x2, y1, y2, xa1, xa2,
                     ya1, ya2, xb1, xb2, yb1, yb2, f1, f2);
        double g1 = -v + vprev;
        double g2 = v - vl;
        g1 = g1 + cf * f1;
        g2 = g2 + cf * f2;
        if (facc) {
            g1 = g1 + facc[s];
            g2 = g2 + facc[V + s];
        }
        mom[s] = m1 + eps * g1;
        mom[V + s] = m2 + eps * g2;
        vprev = v;
        u1 = u1n;
        u2 = u2n;
        x1 = xb1;
        x2 = xb2;
        y1 = yb1;
        y2 = yb2;
    }
}

// sin x and cos x.  For |x| <= pi/4, the range every step eps p of an
// integrator falls in, the Taylor series to x^19 and x^20 (Horner, FMA): the
// first omitted terms are below 1e-21 relative there, so the result is the
// rounding of the series sums (a few ulp at most); sincos() otherwise.
__device__ __forceinline__ void link_sincos(double x, double *s, double *c)
{
    if (!(fabs(x) <= 0.78539816339744830962)) {
        sincos(x, s, c);
        return;
    }
    const double z = x * x;
    // sin x = x + x z (-1/3! + z/5! - ... - z^8/19!)
    double ps = -1.0 / 121645100408832000.0;
    ps = fma(ps, z, 1.0 / 355687428096000.0);
    ps = fma(ps, z, -1.0 / 1307674368000.0);
    ps = fma(ps, z, 1.0 / 6227020800.0);
    ps = fma(ps, z, -1.0 / 39916800.0);
    ps = fma(ps, z, 1.0 / 362880.0);
    ps = fma(ps, z, -1.0 / 5040.0);
    ps = fma(ps, z, 1.0 / 120.0);
    ps = fma(ps, z, -1.0 / 6.0);
    *s = fma(x * z, ps, x);
    // cos x = 1 + z (-1/2! + z/4! - ... + z^9/20!)
    double pc = 1.0 / 2432902008176640000.0;
    pc = fma(pc, z, -1.0 / 6402373705728000.0);
    pc = fma(pc, z, 1.0 / 20922789888000.0);
    pc = fma(pc, z, -1.0 / 87178291200.0);
    pc = fma(pc, z, 1.0 / 479001600.0);
    pc = fma(pc, z, -1.0 / 3628800.0);
    pc = fma(pc, z, 1.0 / 40320.0);
    pc = fma(pc, z, -1.0 / 720.0);
    pc = fma(pc, z, 1.0 / 24.0);
    pc = fma(pc, z, -0.5);
    *c = fma(z, pc, 1.0);
}

// exp(i eps m) u, the link update (src/QED2dAction.jl:55-64).
__device__ __forceinline__ double2 updt_link(double2 u, double m, double eps)
{
    double sn, cs;
    link_sincos(eps * m, &sn, &cs);
    return cmul<double>(mk<double>(cs, sn), u);
}

// The link update U(eu) and then the pure-gauge momentum update P(ep) of an
// integrator stage pair, in one pass: U' = exp(i eu p) U, p' = p + ep F(U').
// Out of place (Ui, mi -> Uo, mo), since the tiles around a tile read the old
// links and momenta of its edge rows.  A march as k_gauge_mom's, except that
// the waves overlap: a wave covers 64 consecutive columns (modulo L1) and
// stores the 62 inside ones, so the updated U2(x+1) and the plaquette of
// x - 1 always come from an adjacent lane and no lane forms extra updated
// links (each an FP64 sincos) while the rest of its wave waits.  Each lane
// updates its own column's links of rows r0 - 1 .. r0 + R.
constexpr int kGaugeOut = kWave - 2;  // columns a wave stores

template <int R>
__global__ void __launch_bounds__(kBlock) k_updt_gauge_mom(Latt L, const double2 *__restrict__ Ui,
                                                           const double *__restrict__ mi,
                                                           double2 *__restrict__ Uo,
                                                           double *__restrict__ mo, double eu,
                                                           double beta, double ep, int nwx,
                                                           int nwaves)
{
    const int g = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;  // wave
    if (g >= nwaves) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int wx = g % nwx, r0 = (g / nwx) * R;
    const int cu = wx * kGaugeOut + lane - 1;  // column before wrapping
    const int i1 = ((cu % L1) + L1) % L1;
    const bool out = (lane >= 1) && (lane <= kGaugeOut) && (cu < L1);
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    const int nr = (r0 + R <= L2) ? R : L2 - r0;
    const long sm = (long)rm * L1, s0 = (long)r0 * L1;
    auto upd = [&](long i) { return updt_link(Ui[i], mi[i], eu); };
    double2 u1 = upd(sm + i1), u2 = upd(V + sm + i1);
    double2 u1n = upd(s0 + i1);
    double2 u2r = from_right<double>(u2);
    double vprev = beta * plaq4(u1, u2r, u1n, u2).y;
    u1 = u1n;
    u2 = upd(V + s0 + i1);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k >= nr) break;
        const int r = r0 + k;
        const int rn = (r + 1 == L2) ? 0 : r + 1;
        const long s = (long)r * L1 + i1, sn = (long)rn * L1;
        u1n = upd(sn + i1);
        double2 u2n{};
        if (k + 1 < nr) u2n = upd(V + sn + i1);
        const double m1 = mi[s], m2 = mi[V + s];
        u2r = from_right<double>(u2);
        const double v = beta * plaq4(u1, u2r, u1n, u2).y;
        const double vl = lane_shift(v, true);
        if (out) {
            Uo[s] = u1;
            Uo[V + s] = u2;
            const double g1 = -v + vprev;
            const double g2 = v - vl;
            mo[s] = m1 + ep * g1;
            mo[V + s] = m2 + ep * g2;
        }
        vprev = v;
        u1 = u1n;
        u2 = u2n;
    }
}

static int gauge_rows(const Latt &L)
{
    int rows = 8;
    while (rows > 1 && march_grid(L, rows).nblocks() < 1024) rows >>= 1;
    return rows;
}

void launch_updt_gauge_mom(const Latt &L, const double2 *Ui, const double *mi, double2 *Uo,
                           double *mo, double eu, double beta, double ep, hipStream_t st)
{
    // (16 rows where the lattice leaves 1024 blocks: each tile recomputes the
    // updated links of its halo rows)
    int rows = 16;
    while (rows > 1 && (long)((L.L1 + kGaugeOut - 1) / kGaugeOut) * ((L.L2 + rows - 1) / rows) <
                           4L * 1024)
        rows >>= 1;
    const int nwx = (L.L1 + kGaugeOut - 1) / kGaugeOut;
    const int nwaves = nwx * ((L.L2 + rows - 1) / rows);
    const int wpb = kBlock / kWave;
    const int nb = (nwaves + wpb - 1) / wpb;
    DISPATCH_ROWS(rows, k_updt_gauge_mom, dim3(nb), dim3(kBlock), 0, st, L, Ui, mi, Uo, mo, eu,
                  beta, ep, nwx, nwaves);
}

void launch_force_mom(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                      const double *facc, double *mom, double beta, double eps, double cf,
                      hipStream_t st)
{
    if (!X) {
        // pure gauge (+ an accumulated force): the marching kernel, 8 rows per
        // thread where the lattice leaves 1024 blocks, fewer below
        const int rows = gauge_rows(L);
        const MarchGrid g = march_grid(L, rows);
        DISPATCH_ROWS(rows, k_gauge_mom, dim3(g.nblocks()), dim3(kBlock), 0, st, L, U, facc,
                      mom, beta, eps, g.bx, g.by, g.nbx);
        return;
    }
    // with the fermion force: 8 rows where the lattice leaves a block per CU
    // (the halo rows every tile re-reads are 1/rows of its X, Y and link
    // loads, mostly from L2; at 1024^2, 16 rows leave one block per CU and ran
    // 30.4 us, 8 rows 25.9-27.3 us, 4 rows 28.6-30.1 us)
    int rows = 8;
    while (rows > 1 && march_grid(L, rows).nblocks() < 256) rows >>= 1;
    const MarchGrid g = march_grid(L, rows);
    DISPATCH_ROWS(rows, k_force_mom, dim3(g.nblocks()), dim3(kBlock), 0, st, L, U, X, Y, facc,
                  mom, beta, eps, cf, g.bx, g.by, g.nbx);
}

// Q X_n for a table of fields (the pole solutions of every pseudofermion of an
// RHMC force): grid y = n, the plain apply's march and stores, mass2[n] the
// field's 2 + m0.
template <int R>
__global__ void __launch_bounds__(kBlock) k_q_apply_tab(Latt L, const double2 *__restrict__ U,
                                                        const double2 *const *X,
                                                        double2 *const *Y, const double *mass2,
                                                        int bx, int by, int nbx)
{
    const int n = blockIdx.y;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    SrcPlain<double> src{X[n], L.V};
    EpiStore<double> epi{Y[n], L.V};
    q_march<double, R>(L, U, mass2[n], src, epi, p.i1, p.r0);
}
void launch_q_apply_tab(const Latt &L, const double2 *U, const double2 *const *X,
                        double2 *const *Y, const double *mass2, int n, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    DISPATCH_ROWS(rows, k_q_apply_tab, dim3(g.nblocks(), n), dim3(kBlock), 0, st, L, U, X, Y,
                  mass2, g.bx, g.by, g.nbx);
}

// The fermion force of rows r0 .. r0+nr-1 of column i1, marching as
// k_force_mom: the links and the X, Y of row r + 1 (x + 2) are loaded one row
// ahead, X, Y of x + 1 come from the adjacent lane (loaded at the wave's right
// edge and at the periodic wrap).  emit(k, s, f1, f2) receives row r0 + k.
template <int R, class Emit>
__device__ __forceinline__ void ferm_march(const Latt &L, const double2 *__restrict__ U,
                                           const double2 *__restrict__ X,
                                           const double2 *__restrict__ Y, int i1, int r0,
                                           int nr, int i1p, bool ldR, Emit emit)
{
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const bool apbc = (L.bc & 1) != 0;
    const long s0 = (long)r0 * L1;
    double2 x1 = X[s0 + i1], x2 = X[V + s0 + i1], y1 = Y[s0 + i1], y2 = Y[V + s0 + i1];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k >= nr) break;
        const int r = r0 + k;
        const int rn = (r + 1 == L2) ? 0 : r + 1;
        const long row = (long)r * L1, s = row + i1, sn = (long)rn * L1 + i1;
        const double2 u1 = U[s], u2 = U[V + s];
        const double2 xb1 = X[sn], xb2 = X[V + sn], yb1 = Y[sn], yb2 = Y[V + sn];
        double2 xa1 = from_right<double>(x1), xa2 = from_right<double>(x2);
        double2 ya1 = from_right<double>(y1), ya2 = from_right<double>(y2);
        if (ldR) {
            xa1 = X[row + i1p];
            xa2 = X[V + row + i1p];
            ya1 = Y[row + i1p];
            ya2 = Y[V + row + i1p];
        }
        double f1, f2;
        ferm_force_v(u1, (apbc && r == L2 - 1) ? cneg<double>(u2) : u2, x1, x2, y1, y2, xa1, xa2,
                     ya1, ya2, xb1, xb2, yb1, yb2, f1, f2);
        emit(k, s, f1, f2);
        x1 = xb1;
        x2 = xb2;
        y1 = yb1;
        y2 = yb2;
    }
}

// facc = sum_n rho[n] F(X_n, Y_n) over a table of fields, per site in the order
// n = 0, 1, ... with the expressions of k_force_acc (the same bits as its
// launches one field after another).  Marches rows as ferm_march, the field
// loop outside the rows, the partial sums of the thread's R sites in registers.
template <int R>
__global__ void __launch_bounds__(kBlock) k_force_acc_tab(Latt L, const double2 *__restrict__ U,
                                                          const double2 *const *X,
                                                          const double2 *const *Y,
                                                          const double *rho, int n,
                                                          double *__restrict__ facc, int bx,
                                                          int by, int nbx)
{
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    // a wave lies in one row of the block (bx is a multiple of the wave size),
    // so r0 is uniform: row offsets go to scalar registers
    const int i1 = p.i1, r0 = __builtin_amdgcn_readfirstlane(p.r0);
    const int i1p = (i1 + 1 == L.L1) ? 0 : i1 + 1;
    const bool ldR = ((threadIdx.x & (kWave - 1)) == kWave - 1) || (i1p == 0);
    const int nr = (r0 + R <= L.L2) ? R : L.L2 - r0;
    // 0 + rho f rounds as rho f, so the first field needs no branch of its own
    double a1[R], a2[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a1[j] = a2[j] = 0.0;
#pragma unroll 1
    for (int k = 0; k < n; ++k) {
        const double rk = rho[k];
        ferm_march<R>(L, U, X[k], Y[k], i1, r0, nr, i1p, ldR,
                      [&](int j, long, double f1, double f2) {
                          a1[j] = a1[j] + rk * f1;
                          a2[j] = a2[j] + rk * f2;
                      });
    }
    const long V = L.V;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        if (j >= nr) break;
        const long s = (long)(r0 + j) * L.L1 + i1;
        facc[s] = a1[j];
        facc[V + s] = a2[j];
    }
}
void launch_force_acc_tab(const Latt &L, const double2 *U, const double2 *const *X,
                          const double2 *const *Y, const double *rho, int n, double *facc,
                          hipStream_t st)
{
    // at most 2 rows: the links and partial sums of every row stay in registers
    // across the fields
    const int rows = gauge_rows(L) < 2 ? gauge_rows(L) : 2;
    const MarchGrid g = march_grid(L, rows);
    DISPATCH_ROWS(rows, k_force_acc_tab, dim3(g.nblocks()), dim3(kBlock), 0, st, L, U, X, Y, rho,
                  n, facc, g.bx, g.by, g.nbx);
}

// facc (+)= rho * F_fermion  (src/QED2dHMC.jl:541-543), marching as ferm_march
template <int R>
__global__ void __launch_bounds__(kBlock) k_force_acc(Latt L, const double2 *__restrict__ U,
                                                      const double2 *__restrict__ X,
                                                      const double2 *__restrict__ Y,
                                                      double *__restrict__ facc, double rho,
                                                      int first, int bx, int by, int nbx)
{
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    const int i1 = p.i1, r0 = p.r0;
    const int i1p = (i1 + 1 == L.L1) ? 0 : i1 + 1;
    const bool ldR = ((threadIdx.x & (kWave - 1)) == kWave - 1) || (i1p == 0);
    const int nr = (r0 + R <= L.L2) ? R : L.L2 - r0;
    const long V = L.V;
    ferm_march<R>(L, U, X, Y, i1, r0, nr, i1p, ldR, [&](int, long s, double f1, double f2) {
        if (first) {
            facc[s] = rho * f1;
            facc[V + s] = rho * f2;
        } else {
            facc[s] = facc[s] + rho * f1;
            facc[V + s] = facc[V + s] + rho * f2;
        }
    });
}
void launch_force_acc(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                      double *facc, double rho, int first, hipStream_t st)
{
    const int rows = gauge_rows(L);
    const MarchGrid g = march_grid(L, rows);
    DISPATCH_ROWS(rows, k_force_acc, dim3(g.nblocks()), dim3(kBlock), 0, st, L, U, X, Y, facc,
                  rho, first, g.bx, g.by, g.nbx);
}

// U <- (cos(eps p), sin(eps p)) U  (src/QED2dAction.jl:55-64)
__global__ void k_updt_fld(double2 *__restrict__ U, const double *__restrict__ mom, double eps,
                           long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        U[i] = updt_link(U[i], mom[i], eps);
    }
}
void launch_updt_fld(const Latt &L, double2 *U, const double *mom, double eps, hipStream_t st)
{
    launch_k(k_updt_fld, dim3(ew_blocks(2 * L.V)), dim3(kBlock), 0, st, U, mom, eps,
                       2 * L.V);
}

// ===========================================================================
// Random fields (counter-based; same streams as the CPU oracle)
// ===========================================================================
__global__ void k_fill_momenta(double *__restrict__ mom, long V, uint32_t k0, uint32_t k1,
                               uint32_t traj)
{
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < V;
         s += (long)gridDim.x * blockDim.x) {
        double g1, g2;
        gauss_pair((uint32_t)s, 0u, kPurposeMom, traj, k0, k1, &g1, &g2);
        mom[s] = g1;
        mom[V + s] = g2;
    }
}
void launch_fill_momenta(const Latt &L, double *mom, uint32_t k0, uint32_t k1, uint32_t traj,
                         hipStream_t st)
{
    launch_k(k_fill_momenta, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, mom, L.V, k0,
                       k1, traj);
}

__global__ void k_fill_cgauss(double2 *__restrict__ X, long V, uint32_t k0, uint32_t k1,
                              uint32_t traj, uint32_t purpose, uint32_t sub0)
{
#pragma clang fp contract(off)
    const double sq2 = 1.4142135623730951;
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < V;
         s += (long)gridDim.x * blockDim.x) {
        for (uint32_t c = 0; c < 2; ++c) {
            double g1, g2;
            gauss_pair((uint32_t)s, sub0 + c, purpose, traj, k0, k1, &g1, &g2);
            X[c * V + s] = mk<double>(g1 / sq2, g2 / sq2);
        }
    }
}
void launch_fill_cgauss(const Latt &L, double2 *X, uint32_t k0, uint32_t k1, uint32_t traj,
                        uint32_t purpose, uint32_t sub0, hipStream_t st)
{
    launch_k(k_fill_cgauss, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, X, L.V, k0, k1,
                       traj, purpose, sub0);
}
void launch_fill_pseudofermion(const Latt &L, double2 *X, uint32_t k0, uint32_t k1,
                               uint32_t traj, uint32_t flavour, hipStream_t st)
{
    launch_fill_cgauss(L, X, k0, k1, traj, kPurposePF + flavour, 0, st);
}

__global__ void k_hot_start(double2 *__restrict__ U, long V, uint32_t k0, uint32_t k1)
{
#pragma clang fp contract(off)
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < V;
         s += (long)gridDim.x * blockDim.x) {
        for (uint32_t d = 0; d < 2; ++d) {
            uint32_t c[4] = {(uint32_t)s, d, kPurposeHot, 0u};
            philox4x32_10(c, k0, k1);
            double u = u53(c[0], c[1]);
            double sn, cs;
            det_sincos2pi(u - 0.5, &sn, &cs);
            U[d * V + s] = mk<double>(cs, sn);
        }
    }
}
void launch_hot_start(const Latt &L, double2 *U, uint32_t k0, uint32_t k1, hipStream_t st)
{
    launch_k(k_hot_start, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, U, L.V, k0, k1);
}

// Lattice doubling (src/QED2dFields.jl:21-50).
__global__ void k_unfold(Latt L, const double2 *__restrict__ Uin, double2 *__restrict__ Uout,
                         int reflect)
{
    const long V2 = 4 * L.V;
    const int M1 = 2 * L.L1;
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)gridDim.x * blockDim.x) {
        const int i2 = (int)(s / L.L1), i1 = (int)(s - (long)i2 * L.L1);
        for (int d = 0; d < 2; ++d) {
            double2 u = Uin[d * L.V + s];
            double2 uc = reflect ? cconj<double>(u) : u;
            Uout[d * V2 + (long)i2 * M1 + i1] = u;
            Uout[d * V2 + (long)i2 * M1 + i1 + L.L1] = uc;
            Uout[d * V2 + (long)(i2 + L.L2) * M1 + i1] = uc;
            Uout[d * V2 + (long)(i2 + L.L2) * M1 + i1 + L.L1] = u;
        }
    }
}
void launch_unfold(const Latt &L, const double2 *Uin, double2 *Uout, int reflect, hipStream_t st)
{
    launch_k(k_unfold, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, Uin, Uout,
                       reflect);
}

}  // namespace qed2d
