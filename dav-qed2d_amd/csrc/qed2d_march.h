// qed2d_march.h -- the marching Wilson stencil shared by every kernel that
// applies Q = gamma5 Dw (plain apply, CG and multi-shift iterations, residuals).
//
// A block holds kBlock threads arranged bx (along i1) x by (along i2).  Each
// thread owns one i1 column and marches up R consecutive rows, keeping the
// spinor rows i2-1, i2, i2+1 and the link U2(x-2) in registers, so each new
// row costs one spinor load, one link load and the epilogue's stores.
#pragma once

#include "qed2d_internal.h"

namespace qed2d {

struct MarchPos {
    int i1, r0;
    bool active;
    int blk;  // logical block index (partial-sum slot)
};

__device__ __forceinline__ MarchPos march_pos(const Latt &L, int bx, int by, int nbx, int rows)
{
    MarchPos p;
    const int nb = gridDim.x;
    const int b = xcd_remap(blockIdx.x, nb);
    const int bxi = b % nbx, byi = b / nbx;
    const int tx = threadIdx.x % bx, ty = threadIdx.x / bx;
    p.i1 = bxi * bx + tx;
    p.r0 = (byi * by + ty) * rows;
    p.active = (p.i1 < L.L1) && (p.r0 < L.L2);
    p.blk = b;
    return p;
}

// Spinor sources for the stencil.
template <typename T> struct SrcPlain {
    const c2<T> *__restrict__ s;
    long V;
    __device__ __forceinline__ void load(long i, c2<T> &a, c2<T> &b) const
    {
        a = s[i];
        b = s[V + i];
    }
};

// p_new = r + beta * p_old, evaluated where it is read (src/QED2dSolver.jl:35).
template <typename T> struct SrcXpay {
    const c2<T> *__restrict__ r;
    const c2<T> *__restrict__ p;
    T beta;
    long V;
    __device__ __forceinline__ void load(long i, c2<T> &a, c2<T> &b) const
    {
        c2<T> r0 = r[i], r1 = r[V + i], p0 = p[i], p1 = p[V + i];
        a = mk<T>(r0.x + beta * p0.x, r0.y + beta * p0.y);
        b = mk<T>(r1.x + beta * p1.x, r1.y + beta * p1.y);
    }
};

// March one thread's column: for rows r0 .. r0+R-1 compute (Q s)(x) and hand
// it to the epilogue together with s(x).
template <typename T, int R, typename Src, typename Epi>
__device__ __forceinline__ void q_march(const Latt &L, const c2<T> *__restrict__ U, T mass2,
                                        const Src &src, Epi &epi, int i1, int r0)
{
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int i1p = (i1 + 1 == L1) ? 0 : i1 + 1;
    const int i1m = (i1 == 0) ? L1 - 1 : i1 - 1;
    const bool ap = (L.bc & 1) != 0;
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    c2<T> dn1, dn2, c1, cc2;
    src.load((long)rm * L1 + i1, dn1, dn2);
    src.load((long)r0 * L1 + i1, c1, cc2);
    c2<T> U2m = U[V + (long)rm * L1 + i1];
    if (ap && r0 == 0) U2m = cneg<T>(U2m);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = r0 + k;
        if (r >= L2) break;
        const int rp = (r + 1 == L2) ? 0 : r + 1;
        const long row = (long)r * L1;
        const long s = row + i1;
        c2<T> up1, up2, p11, p12, m11, m12;
        src.load((long)rp * L1 + i1, up1, up2);
        src.load(row + i1p, p11, p12);
        src.load(row + i1m, m11, m12);
        const c2<T> U1 = U[s], U1m = U[row + i1m], U2 = U[V + s];
        const c2<T> U2e = (ap && r == L2 - 1) ? cneg<T>(U2) : U2;
        c2<T> o1, o2;
        q_site<T>(mass2, c1, cc2, p11, p12, m11, m12, up1, up2, dn1, dn2, U1, U1m, U2e, U2m, o1,
                  o2);
        epi(s, o1, o2, c1, cc2);
        dn1 = c1;
        dn2 = cc2;
        c1 = up1;
        cc2 = up2;
        U2m = U2;
    }
}


#define DISPATCH_ROWS(rows, KERNEL, ...)                                              \
    switch (rows) {                                                                   \
    case 1: hipLaunchKernelGGL((KERNEL<1>), __VA_ARGS__); break;                      \
    case 2: hipLaunchKernelGGL((KERNEL<2>), __VA_ARGS__); break;                      \
    case 4: hipLaunchKernelGGL((KERNEL<4>), __VA_ARGS__); break;                      \
    case 8: hipLaunchKernelGGL((KERNEL<8>), __VA_ARGS__); break;                      \
    default: hipLaunchKernelGGL((KERNEL<16>), __VA_ARGS__); break;                    \
    }

// Sum of v over the block, returned to every thread.  lds: kBlock/64 + 1 doubles.
__device__ __forceinline__ double block_allsum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += lds[i];
        lds[nw] = s;
    }
    __syncthreads();
    const double r = lds[nw];
    __syncthreads();
    return r;
}

// Sum of the per-block partials part[b * stride + off], b < n, in a fixed
// order; every block that calls this gets the same bits.  The partials were
// written by an earlier kernel, so the kernel boundary has made them visible.
__device__ __forceinline__ double sum_prev_partials(const double *part, int n, int stride,
                                                    int off, double *lds)
{
    double s = 0.0;
    for (int b = threadIdx.x; b < n; b += blockDim.x) s += part[(long)b * stride + off];
    return block_allsum(s, lds);
}

}  // namespace qed2d
