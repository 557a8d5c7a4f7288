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


// ===========================================================================
// FP32 with two sites per thread along i1: a thread owns the column pair
// (2j, 2j+1), so every spinor and link access is a 16-byte float4 as in the
// FP64 kernels.  Needs L1 even and L1 >= 4.
// ===========================================================================
struct F2x2 {
    float2 a0, a1;  // sites 2j and 2j+1 of one component
};

__device__ __forceinline__ F2x2 ld2(const float2 *__restrict__ p, long i)
{
    const float4 v = *reinterpret_cast<const float4 *>(p + i);
    F2x2 r;
    r.a0 = make_float2(v.x, v.y);
    r.a1 = make_float2(v.z, v.w);
    return r;
}
__device__ __forceinline__ void st2(float2 *__restrict__ p, long i, float2 a, float2 b)
{
    *reinterpret_cast<float4 *>(p + i) = make_float4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ F2x2 neg2(F2x2 v)
{
    v.a0 = cneg<float>(v.a0);
    v.a1 = cneg<float>(v.a1);
    return v;
}

__device__ __forceinline__ MarchPos march_pos2(const Latt &L, int bx, int by, int nbx, int rows)
{
    MarchPos p;
    const int nb = gridDim.x;
    const int b = xcd_remap(blockIdx.x, nb);
    const int bxi = b % nbx, byi = b / nbx;
    const int tx = threadIdx.x % bx, ty = threadIdx.x / bx;
    p.i1 = bxi * bx + tx;  // pair index j
    p.r0 = (byi * by + ty) * rows;
    p.active = (p.i1 < L.L1 / 2) && (p.r0 < L.L2);
    p.blk = b;
    return p;
}

struct SrcPlain2 {
    const float2 *__restrict__ s;
    long V;
    __device__ __forceinline__ void load2(long i, F2x2 &a, F2x2 &b) const
    {
        a = ld2(s, i);
        b = ld2(s, V + i);
    }
    __device__ __forceinline__ void load1(long i, float2 &a, float2 &b) const
    {
        a = s[i];
        b = s[V + i];
    }
};

struct SrcXpay2 {
    const float2 *__restrict__ r;
    const float2 *__restrict__ p;
    float beta;
    long V;
    __device__ __forceinline__ float2 xp(float2 a, float2 b) const
    {
        return make_float2(a.x + beta * b.x, a.y + beta * b.y);
    }
    __device__ __forceinline__ void load2(long i, F2x2 &a, F2x2 &b) const
    {
        const F2x2 r0 = ld2(r, i), r1 = ld2(r, V + i), p0 = ld2(p, i), p1 = ld2(p, V + i);
        a.a0 = xp(r0.a0, p0.a0);
        a.a1 = xp(r0.a1, p0.a1);
        b.a0 = xp(r1.a0, p1.a0);
        b.a1 = xp(r1.a1, p1.a1);
    }
    __device__ __forceinline__ void load1(long i, float2 &a, float2 &b) const
    {
        a = xp(r[i], p[i]);
        b = xp(r[V + i], p[V + i]);
    }
};

// March the column pair (2j, 2j+1) over rows r0 .. r0+R-1; the epilogue gets
// both sites' outputs and inputs: epi.pair(s, oa1, oa2, ob1, ob2, c1, c2) with
// s the index of site 2j, (oa, ob) the outputs of sites 2j, 2j+1 for spin
// components 1, 2 and c1, c2 the inputs.
template <int R, typename Src, typename Epi>
__device__ __forceinline__ void q_march2(const Latt &L, const float2 *__restrict__ U, float mass2,
                                         const Src &src, Epi &epi, int j, int r0)
{
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int ia = 2 * j;
    const int im = (ia == 0) ? L1 - 1 : ia - 1;
    const int ip = (ia + 2 == L1) ? 0 : ia + 2;
    const bool ap = (L.bc & 1) != 0;
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    F2x2 dn1, dn2, c1, cc2;
    src.load2((long)rm * L1 + ia, dn1, dn2);
    src.load2((long)r0 * L1 + ia, c1, cc2);
    F2x2 U2m = ld2(U, V + (long)rm * L1 + ia);
    if (ap && r0 == 0) U2m = neg2(U2m);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = r0 + k;
        if (r >= L2) break;
        const int rp = (r + 1 == L2) ? 0 : r + 1;
        const long row = (long)r * L1;
        const long s = row + ia;
        F2x2 up1, up2;
        float2 l1, l2, q1, q2;
        src.load2((long)rp * L1 + ia, up1, up2);
        src.load1(row + im, l1, l2);
        src.load1(row + ip, q1, q2);
        const F2x2 U1 = ld2(U, s), U2 = ld2(U, V + s);
        const float2 U1l = U[row + im];
        const F2x2 U2e = (ap && r == L2 - 1) ? neg2(U2) : U2;
        float2 oa1, oa2, ob1, ob2;
        // site 2j: x+1 = 2j+1, x-1 = the left neighbour
        q_site<float>(mass2, c1.a0, cc2.a0, c1.a1, cc2.a1, l1, l2, up1.a0, up2.a0, dn1.a0,
                      dn2.a0, U1.a0, U1l, U2e.a0, U2m.a0, oa1, oa2);
        // site 2j+1: x+1 = the right neighbour, x-1 = 2j
        q_site<float>(mass2, c1.a1, cc2.a1, q1, q2, c1.a0, cc2.a0, up1.a1, up2.a1, dn1.a1,
                      dn2.a1, U1.a1, U1.a0, U2e.a1, U2m.a1, ob1, ob2);
        epi.pair(s, oa1, oa2, ob1, ob2, c1, cc2);
        dn1 = c1;
        dn2 = cc2;
        c1 = up1;
        cc2 = up2;
        U2m = U2;
    }
}

// Launch shape of the pair kernels (width L1/2).
inline MarchGrid march_grid2(const Latt &L, int rows)
{
    Latt h = L;
    h.L1 = L.L1 / 2;
    return march_grid(h, rows);
}

inline bool pair_ok(const Latt &L) { return (L.L1 % 2) == 0 && L.L1 >= 4; }

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
