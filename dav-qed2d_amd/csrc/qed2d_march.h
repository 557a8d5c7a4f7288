// qed2d_march.h -- the marching Wilson stencil shared by every kernel that
// applies Q = gamma5 Dw (plain apply, CG and multi-shift iterations, residuals).
//
// A block holds kBlock threads arranged bx (along i1) x by (along i2).  Each
// thread owns one i1 column and marches up R consecutive rows, keeping the
// spinor rows i2-1, i2, i2+1 and the link U2(x-2) in registers, so each new
// row costs one spinor load, one link load and the epilogue's stores.  The
// march is software pipelined one row deep: a wave keeps two rows of loads in
// flight, which is what hides HBM latency at the one or two waves per SIMD a
// marching launch leaves on a 1024^2 lattice.
#pragma once

#include "qed2d_internal.h"

namespace qed2d {

struct MarchPos {
    int i1, r0;
    bool active;
    int blk;  // logical block index (partial-sum slot)
};

// Block order of the plain Q applies.  XCD bands (xcd_remap) keep the
// vertical neighbours of a tile in one L2, which pays while the apply's
// operands are largely served from the caches; from 4096^2 sites on, launch
// order runs faster (same box, best tile of 2, 4, 8, TB/s, bands -> launch
// order: 4096^2 FP64 5.89 -> 6.00, FP32 5.70 -> 5.98; 6144^2 5.76 -> 5.97,
// 5.56 -> 5.68; 8192^2 5.80 -> 5.83, 5.79 -> 5.78; at 2048^2 bands win, 7.19
// and 6.87 against 6.81 and 6.02).  The solver kernels keep the bands at every
// size (launch order at 4096^2: mixed Nf=2 trajectory 1.5 % slower).
__device__ __forceinline__ bool apply_bands(const Latt &L) { return L.V < (1L << 24); }

__device__ __forceinline__ MarchPos march_pos(const Latt &L, int bx, int by, int nbx, int rows,
                                              bool bands = true)
{
    MarchPos p;
    const int nb = gridDim.x;
    const int b = bands ? xcd_remap(blockIdx.x, nb) : (int)blockIdx.x;
    const int bxi = b % nbx, byi = b / nbx;
    const int tx = threadIdx.x % bx, ty = threadIdx.x / bx;
    p.i1 = bxi * bx + tx;
    p.r0 = (byi * by + ty) * rows;
    p.active = (p.i1 < L.L1) && (p.r0 < L.L2);
    p.blk = b;
    return p;
}

// Spinor sources for the stencil.  fetch() issues the loads of one site and
// returns them raw; value() turns them into the two spin components.  The
// split lets the march issue the loads of the next row before it computes the
// current one.
template <typename T> struct SrcPlain {
    const c2<T> *__restrict__ s;
    long V;
    struct Raw {
        c2<T> a, b;
    };
    __device__ __forceinline__ Raw fetch(long i) const { return Raw{s[i], s[V + i]}; }
    __device__ __forceinline__ void value(const Raw &w, c2<T> &a, c2<T> &b) const
    {
        a = w.a;
        b = w.b;
    }
};

// p_new = r + beta * p_old, evaluated where it is read (src/QED2dSolver.jl:35).
template <typename T> struct SrcXpay {
    const c2<T> *__restrict__ r;
    const c2<T> *__restrict__ p;
    T beta;
    long V;
    struct Raw {
        c2<T> r0, r1, p0, p1;
    };
    __device__ __forceinline__ Raw fetch(long i) const { return Raw{r[i], r[V + i], p[i], p[V + i]}; }
    __device__ __forceinline__ void value(const Raw &w, c2<T> &a, c2<T> &b) const
    {
        a = mk<T>(w.r0.x + beta * w.p0.x, w.r0.y + beta * w.p0.y);
        b = mk<T>(w.r1.x + beta * w.p1.x, w.r1.y + beta * w.p1.y);
    }
};

// Epilogues: pre(s) issues the loads the epilogue needs at site s (fetched one
// row ahead with the stencil operands); operator()(s, pre, o1, o2, c1, c2)
// receives Q s at x (o1, o2) and s(x) itself (c1, c2).  NoPre is the empty
// pre-load of epilogues that only store.
struct NoPre {};

// Stores; NT selects nontemporal stores (streamed past the caches, for
// outputs that are not read back soon).
template <typename T, bool NT> __device__ __forceinline__ void st(c2<T> *p, long i, c2<T> v)
{
    if constexpr (NT) {
        __builtin_nontemporal_store(v.x, &p[i].x);
        __builtin_nontemporal_store(v.y, &p[i].y);
    } else {
        p[i] = v;
    }
}

// Values of the neighbouring lanes (DPP wavefront shifts): from_left(v) is
// lane i-1's v at lane i (wave_shr:1), from_right(v) lane i+1's (wave_shl:1).
// A wave covers 64 consecutive sites along i1, so these are the i1 -/+ 1
// neighbours except at the wave's edges and at the periodic wrap, where the
// march loads them instead.  Every lane executes the shift; lanes whose
// source is out of range receive 0 and load.
__device__ __forceinline__ float lane_shift(float v, bool from_left)
{
    const int i = __float_as_int(v);
    const int r = from_left ? __builtin_amdgcn_update_dpp(0, i, 0x138, 0xF, 0xF, false)
                            : __builtin_amdgcn_update_dpp(0, i, 0x130, 0xF, 0xF, false);
    return __int_as_float(r);
}
__device__ __forceinline__ double lane_shift(double v, bool from_left)
{
    const long long b = __double_as_longlong(v);
    const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    const int rlo = from_left ? __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, false)
                              : __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, false);
    const int rhi = from_left ? __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, false)
                              : __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, false);
    return __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
}
template <typename T> __device__ __forceinline__ c2<T> from_left(c2<T> v)
{
    return mk<T>(lane_shift(v.x, true), lane_shift(v.y, true));
}
template <typename T> __device__ __forceinline__ c2<T> from_right(c2<T> v)
{
    return mk<T>(lane_shift(v.x, false), lane_shift(v.y, false));
}

// March one thread's column over rows r0 .. r0+R-1: compute (Q s)(x) and hand
// it to the epilogue together with s(x).  The operands of row r+1 (the spinor
// of row r+2, the row's links, the epilogue's pre-loads, and the i1
// neighbours of the wave's edge lanes) are in flight while row r is computed;
// the i1 neighbours and U(x-1,1) of the other lanes come from the adjacent
// lanes' registers.
//
// ready() runs after the loads of the first rows are issued and before any of
// them is used (see q_march2r); active = false skips the loads and the rows
// but still runs ready().
template <typename T, int R, typename Src, typename Epi, typename Ready>
__device__ __forceinline__ bool q_march_r(const Latt &L, const c2<T> *__restrict__ U, T mass2,
                                          Src &src, Epi &epi, int i1, int r0, bool active,
                                          Ready &&ready)
{
    using Raw = typename Src::Raw;
    using Pre = decltype(epi.pre(0L));
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int i1p = (i1 + 1 == L1) ? 0 : i1 + 1;
    const int i1m = (i1 == 0) ? L1 - 1 : i1 - 1;
    const int lane = threadIdx.x & (kWave - 1);
    const bool ldR = (lane == kWave - 1) || (i1p == 0);  // i1+1 not in lane+1
    const bool ldL = (lane == 0) || (i1 == 0);           // i1-1 not in lane-1
    const bool ap = (L.bc & 1) != 0;
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    const int nr = (r0 + R <= L2) ? R : L2 - r0;
    // operands of rows r0 - 1, r0 and r0 + 1
    const long row = (long)r0 * L1;
    Raw wd{}, wc{}, wu{}, wp{}, wm{};
    c2<T> U2m{}, U1m{}, U1{}, U2{};
    Pre q{};
    if (active) {
        wd = src.fetch((long)rm * L1 + i1);
        wc = src.fetch(row + i1);
        U2m = U[V + (long)rm * L1 + i1];
        wu = src.fetch((long)((r0 + 1 == L2) ? 0 : r0 + 1) * L1 + i1);
        if (ldR) wp = src.fetch(row + i1p);
        if (ldL) {
            wm = src.fetch(row + i1m);
            U1m = U[row + i1m];
        }
        U1 = U[row + i1];
        U2 = U[V + row + i1];
        q = epi.pre(row + i1);
    }
    if (!ready()) return false;
    if (!active) return true;
    c2<T> dn1, dn2, c1, cc2;
    src.value(wd, dn1, dn2);
    src.value(wc, c1, cc2);
    if (ap && r0 == 0) U2m = cneg<T>(U2m);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k >= nr) break;
        const int r = r0 + k;
        const bool more = (k + 1 < nr);
        Raw nwu, nwp{}, nwm{};
        c2<T> nU1, nU1m{}, nU2;
        Pre nq;
        if (more) {
            const long row1 = (long)(r + 1) * L1;
            const int r2 = (r + 2 >= L2) ? r + 2 - L2 : r + 2;
            nwu = src.fetch((long)r2 * L1 + i1);
            if (ldR) nwp = src.fetch(row1 + i1p);
            if (ldL) {
                nwm = src.fetch(row1 + i1m);
                nU1m = U[row1 + i1m];
            }
            nU1 = U[row1 + i1];
            nU2 = U[V + row1 + i1];
            nq = epi.pre(row1 + i1);
        }
        c2<T> up1, up2, p11, p12, m11, m12;
        src.value(wu, up1, up2);
        // neighbours along i1 from the adjacent lanes (every lane shifts)
        p11 = from_right<T>(c1);
        p12 = from_right<T>(cc2);
        m11 = from_left<T>(c1);
        m12 = from_left<T>(cc2);
        c2<T> U1l = from_left<T>(U1);
        if (ldR) src.value(wp, p11, p12);
        if (ldL) {
            src.value(wm, m11, m12);
            U1l = U1m;
        }
        const c2<T> U2e = (ap && r == L2 - 1) ? cneg<T>(U2) : U2;
        c2<T> o1, o2;
        q_site<T>(mass2, c1, cc2, p11, p12, m11, m12, up1, up2, dn1, dn2, U1, U1l, U2e, U2m, o1,
                  o2);
        epi((long)r * L1 + i1, q, o1, o2, c1, cc2);
        dn1 = c1;
        dn2 = cc2;
        c1 = up1;
        cc2 = up2;
        U2m = U2;
        if (more) {
            wu = nwu;
            wp = nwp;
            wm = nwm;
            U1 = nU1;
            U1m = nU1m;
            U2 = nU2;
            q = nq;
        }
    }
    return true;
}

// q_march_r without a ready step.
template <typename T, int R, typename Src, typename Epi>
__device__ __forceinline__ void q_march(const Latt &L, const c2<T> *__restrict__ U, T mass2,
                                        const Src &src, Epi &epi, int i1, int r0)
{
    Src s = src;
    q_march_r<T, R>(L, U, mass2, s, epi, i1, r0, true, [] { return true; });
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
template <bool NT = false>
__device__ __forceinline__ void st2(float2 *__restrict__ p, long i, float2 a, float2 b)
{
    if constexpr (NT) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        const v4f v = {a.x, a.y, b.x, b.y};
        __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(p + i));
    } else {
        *reinterpret_cast<float4 *>(p + i) = make_float4(a.x, a.y, b.x, b.y);
    }
}
__device__ __forceinline__ F2x2 neg2(F2x2 v)
{
    v.a0 = cneg<float>(v.a0);
    v.a1 = cneg<float>(v.a1);
    return v;
}

__device__ __forceinline__ MarchPos march_pos2(const Latt &L, int bx, int by, int nbx, int rows,
                                               bool bands = true)
{
    MarchPos p;
    const int nb = gridDim.x;
    const int b = bands ? xcd_remap(blockIdx.x, nb) : (int)blockIdx.x;
    const int bxi = b % nbx, byi = b / nbx;
    const int tx = threadIdx.x % bx, ty = threadIdx.x / bx;
    p.i1 = bxi * bx + tx;  // pair index j
    p.r0 = (byi * by + ty) * rows;
    p.active = (p.i1 < L.L1 / 2) && (p.r0 < L.L2);
    p.blk = b;
    return p;
}

// Pair sources: fetch2/value2 for the column pair, fetch1/value1 for one site.
struct SrcPlain2 {
    const float2 *__restrict__ s;
    long V;
    struct Raw2 {
        F2x2 a, b;
    };
    struct Raw1 {
        float2 a, b;
    };
    __device__ __forceinline__ Raw2 fetch2(long i) const { return Raw2{ld2(s, i), ld2(s, V + i)}; }
    __device__ __forceinline__ Raw1 fetch1(long i) const { return Raw1{s[i], s[V + i]}; }
    __device__ __forceinline__ void value2(const Raw2 &w, F2x2 &a, F2x2 &b) const
    {
        a = w.a;
        b = w.b;
    }
    __device__ __forceinline__ void value1(const Raw1 &w, float2 &a, float2 &b) const
    {
        a = w.a;
        b = w.b;
    }
};

struct SrcXpay2 {
    const float2 *__restrict__ r;
    const float2 *__restrict__ p;
    float beta;
    long V;
    struct Raw2 {
        F2x2 r0, r1, p0, p1;
    };
    struct Raw1 {
        float2 r0, r1, p0, p1;
    };
    __device__ __forceinline__ float2 xp(float2 a, float2 b) const
    {
        return make_float2(a.x + beta * b.x, a.y + beta * b.y);
    }
    __device__ __forceinline__ Raw2 fetch2(long i) const
    {
        return Raw2{ld2(r, i), ld2(r, V + i), ld2(p, i), ld2(p, V + i)};
    }
    __device__ __forceinline__ Raw1 fetch1(long i) const
    {
        return Raw1{r[i], r[V + i], p[i], p[V + i]};
    }
    __device__ __forceinline__ void value2(const Raw2 &w, F2x2 &a, F2x2 &b) const
    {
        a.a0 = xp(w.r0.a0, w.p0.a0);
        a.a1 = xp(w.r0.a1, w.p0.a1);
        b.a0 = xp(w.r1.a0, w.p1.a0);
        b.a1 = xp(w.r1.a1, w.p1.a1);
    }
    __device__ __forceinline__ void value1(const Raw1 &w, float2 &a, float2 &b) const
    {
        a = xp(w.r0, w.p0);
        b = xp(w.r1, w.p1);
    }
};

// The epilogue's loads for the row of site s: epi.pre_raw(s, raw) when the
// epilogue has it (raw: the source's raw loads of that row, already in
// registers), else epi.pre(s).
template <typename E, typename Raw>
__device__ __forceinline__ auto pre_of(const E &e, long s, const Raw &raw, int)
    -> decltype(e.pre_raw(s, raw))
{
    return e.pre_raw(s, raw);
}
template <typename E, typename Raw>
__device__ __forceinline__ auto pre_of(const E &e, long s, const Raw &, long) -> decltype(e.pre(s))
{
    return e.pre(s);
}

// March the column pair (2j, 2j+1) over rows r0 .. r0+R-1, software pipelined
// like q_march, with the neighbours 2j-1 and 2j+2 (and U(2j-1, 1)) from the
// adjacent lanes away from the wave's edges; the epilogue gets both sites'
// outputs and inputs: epi.pair(s, pre, oa1, oa2, ob1, ob2, c1, c2) with s the
// index of site 2j, (oa, ob) the outputs of sites 2j, 2j+1 for spin
// components 1, 2 and c1, c2 the inputs.  The links come through the policy
// lk (LinksF2 or LinksPacked): their raw loads travel one row ahead like the
// spinors' and are decoded where they are used.
//
// ready() runs after the loads of the first rows are issued and before any of
// them is used: a solver kernel finishes there the sums of the previous
// kernel's partials and sets the scalars the source and the epilogue use
// (beta, alpha), so their latency overlaps the first rows' loads.  When it
// returns false the march stops (no output) and returns false.
template <int R, typename Lk, typename Src, typename Epi, typename Ready>
__device__ __forceinline__ bool q_march2r(const Latt &L, const Lk &lk, float mass2, Src &src,
                                          Epi &epi, int j, int r0, bool active, Ready &&ready)
{
    using Raw2 = typename Src::Raw2;
    using Raw1 = typename Src::Raw1;
    using LR2 = typename Lk::Raw2;
    using LR1 = typename Lk::Raw1;
    using Pre = decltype(pre_of(epi, 0L, Raw2{}, 0));
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int ia = 2 * j;
    const int im = (ia == 0) ? L1 - 1 : ia - 1;
    const int ip = (ia + 2 == L1) ? 0 : ia + 2;
    const int lane = threadIdx.x & (kWave - 1);
    const bool ldR = (lane == kWave - 1) || (ip == 0);
    const bool ldL = (lane == 0) || (ia == 0);
    const bool ap = (L.bc & 1) != 0;
    const int rm = (r0 == 0) ? L2 - 1 : r0 - 1;
    const int nr = (r0 + R <= L2) ? R : L2 - r0;
    const long row = (long)r0 * L1;
    Raw2 wd{}, wc{}, wu{};
    Raw1 wl{}, wq{};
    LR2 rU2m{}, rU1{}, rU2{};
    LR1 rU1lm{};
    Pre q{};
    if (active) {
        wd = src.fetch2((long)rm * L1 + ia);
        wc = src.fetch2(row + ia);
        rU2m = lk.fetch2(V + (long)rm * L1 + ia);
        wu = src.fetch2((long)((r0 + 1 == L2) ? 0 : r0 + 1) * L1 + ia);
        if (ldL) {
            wl = src.fetch1(row + im);
            rU1lm = lk.fetch1(row + im);
        }
        if (ldR) wq = src.fetch1(row + ip);
        rU1 = lk.fetch2(row + ia);
        rU2 = lk.fetch2(V + row + ia);
        q = pre_of(epi, row + ia, wc, 0);
    }
    if (!ready()) return false;
    if (!active) return true;
    F2x2 dn1, dn2, c1, cc2;
    src.value2(wd, dn1, dn2);
    src.value2(wc, c1, cc2);
    F2x2 U2m = lk.value2(rU2m);
    if (ap && r0 == 0) U2m = neg2(U2m);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k >= nr) break;
        const int r = r0 + k;
        const bool more = (k + 1 < nr);
        Raw2 nwu;
        Raw1 nwl{}, nwq{};
        LR2 nrU1, nrU2;
        LR1 nrU1lm{};
        Pre nq;
        if (more) {
            const long row1 = (long)(r + 1) * L1;
            const int r2 = (r + 2 >= L2) ? r + 2 - L2 : r + 2;
            nwu = src.fetch2((long)r2 * L1 + ia);
            if (ldL) {
                nwl = src.fetch1(row1 + im);
                nrU1lm = lk.fetch1(row1 + im);
            }
            if (ldR) nwq = src.fetch1(row1 + ip);
            nrU1 = lk.fetch2(row1 + ia);
            nrU2 = lk.fetch2(V + row1 + ia);
            nq = pre_of(epi, row1 + ia, wu, 0);
        }
        const F2x2 U1 = lk.value2(rU1), U2 = lk.value2(rU2);
        F2x2 up1, up2;
        float2 l1, l2, q1, q2;
        src.value2(wu, up1, up2);
        // neighbours 2j-1 (lane-1's second site) and 2j+2 (lane+1's first site)
        l1 = from_left<float>(c1.a1);
        l2 = from_left<float>(cc2.a1);
        q1 = from_right<float>(c1.a0);
        q2 = from_right<float>(cc2.a0);
        float2 U1l = from_left<float>(U1.a1);
        if (ldL) {
            src.value1(wl, l1, l2);
            U1l = lk.value1(rU1lm);
        }
        if (ldR) src.value1(wq, q1, q2);
        const F2x2 U2e = (ap && r == L2 - 1) ? neg2(U2) : U2;
        float2 oa1, oa2, ob1, ob2;
        // site 2j: x+1 = 2j+1, x-1 = the left neighbour
        q_site<float>(mass2, c1.a0, cc2.a0, c1.a1, cc2.a1, l1, l2, up1.a0, up2.a0, dn1.a0,
                      dn2.a0, U1.a0, U1l, U2e.a0, U2m.a0, oa1, oa2);
        // site 2j+1: x+1 = the right neighbour, x-1 = 2j
        q_site<float>(mass2, c1.a1, cc2.a1, q1, q2, c1.a0, cc2.a0, up1.a1, up2.a1, dn1.a1,
                      dn2.a1, U1.a1, U1.a0, U2e.a1, U2m.a1, ob1, ob2);
        epi.pair((long)r * L1 + ia, q, oa1, oa2, ob1, ob2, c1, cc2);
        dn1 = c1;
        dn2 = cc2;
        c1 = up1;
        cc2 = up2;
        U2m = U2;
        if (more) {
            wu = nwu;
            wl = nwl;
            wq = nwq;
            rU1 = nrU1;
            rU2 = nrU2;
            rU1lm = nrU1lm;
            q = nq;
        }
    }
    return true;
}

// q_march2r without a ready step.
template <int R, typename Lk, typename Src, typename Epi>
__device__ __forceinline__ void q_march2l(const Latt &L, const Lk &lk, float mass2, const Src &src,
                                          Epi &epi, int j, int r0)
{
    Src s = src;
    q_march2r<R>(L, lk, mass2, s, epi, j, r0, true, [] { return true; });
}

// FP32 complex links.
struct LinksF2 {
    const float2 *__restrict__ U;
    using Raw2 = F2x2;
    using Raw1 = float2;
    __device__ __forceinline__ F2x2 fetch2(long i) const { return ld2(U, i); }
    __device__ __forceinline__ float2 fetch1(long i) const { return U[i]; }
    __device__ __forceinline__ F2x2 value2(const F2x2 &r) const { return r; }
    __device__ __forceinline__ float2 value1(const float2 &r) const { return r; }
};

// U(1) links packed into one float each (k_link_pack): the smaller of |Re|,
// |Im| with the last mantissa bit cleared; bit 30 (always 0 for magnitudes
// below 2) set when that is Re, the last bit set when the other component is
// negative.  The other component is sqrt(1 - v^2), v^2 <= 1/2.
__device__ __forceinline__ float2 link_dec(float w)
{
    const uint32_t b = __float_as_uint(w);
    const float v = __uint_as_float(b & 0xBFFFFFFEu);
    float o = __builtin_amdgcn_sqrtf(__builtin_fmaf(-v, v, 1.0f));
    if (b & 1u) o = -o;
    return (b & 0x40000000u) ? make_float2(v, o) : make_float2(o, v);
}
struct LinksPacked {
    const float *__restrict__ U;
    using Raw2 = float2;
    using Raw1 = float;
    __device__ __forceinline__ float2 fetch2(long i) const
    {
        return *reinterpret_cast<const float2 *>(U + i);
    }
    __device__ __forceinline__ float fetch1(long i) const { return U[i]; }
    __device__ __forceinline__ F2x2 value2(const float2 &r) const
    {
        F2x2 u;
        u.a0 = link_dec(r.x);
        u.a1 = link_dec(r.y);
        return u;
    }
    __device__ __forceinline__ float2 value1(float r) const { return link_dec(r); }
};

// q_march2l over FP32 links whose format is fixed at compile time (PK: packed
// U(1) links, else complex).
template <int R, bool PK, typename Src, typename Epi>
__device__ __forceinline__ void q_march2xt(const Latt &L, const LinksF32 &u, float mass2,
                                           const Src &src, Epi &epi, int j, int r0)
{
    if constexpr (PK) {
        const LinksPacked lk{(const float *)u.U};
        q_march2l<R>(L, lk, mass2, src, epi, j, r0);
    } else {
        const LinksF2 lk{(const float2 *)u.U};
        q_march2l<R>(L, lk, mass2, src, epi, j, r0);
    }
}

// q_march2r over FP32 links whose format is fixed at compile time, so that no
// branch on it sits between a kernel's first loads and its march.
template <int R, bool PK, typename Src, typename Epi, typename Ready>
__device__ __forceinline__ bool q_march2t(const Latt &L, const LinksF32 &u, float mass2, Src &src,
                                          Epi &epi, int j, int r0, bool active, Ready &&ready)
{
    if constexpr (PK) {
        const LinksPacked lk{(const float *)u.U};
        return q_march2r<R>(L, lk, mass2, src, epi, j, r0, active, ready);
    } else {
        const LinksF2 lk{(const float2 *)u.U};
        return q_march2r<R>(L, lk, mass2, src, epi, j, r0, active, ready);
    }
}

// q_march2l with FP32 complex links.
template <int R, typename Src, typename Epi>
__device__ __forceinline__ void q_march2(const Latt &L, const float2 *__restrict__ U, float mass2,
                                         const Src &src, Epi &epi, int j, int r0)
{
    const LinksF2 lk{U};
    q_march2l<R>(L, lk, mass2, src, epi, j, r0);
}



// The apply's output is not read back by this kernel, so its stores stream
// past the caches (nontemporal).
template <typename T> struct EpiStore {
    c2<T> *__restrict__ so;
    long V;
    __device__ __forceinline__ NoPre pre(long) const { return {}; }
    __device__ __forceinline__ void operator()(long s, NoPre, c2<T> o1, c2<T> o2, c2<T>, c2<T>)
    {
        st<T, true>(so, s, o1);
        st<T, true>(so, V + s, o2);
    }
};

struct EpiStore2 {
    float2 *__restrict__ so;
    long V;
    __device__ __forceinline__ NoPre pre(long) const { return {}; }
    __device__ __forceinline__ void pair(long s, NoPre, float2 oa1, float2 oa2, float2 ob1,
                                         float2 ob2, F2x2, F2x2)
    {
        st2<true>(so, s, oa1, ob1);
        st2<true>(so, V + s, oa2, ob2);
    }
};

// Launch shape of the pair kernels (width L1/2).
inline MarchGrid march_grid2(const Latt &L, int rows)
{
    Latt h = L;
    h.L1 = L.L1 / 2;
    return march_grid(h, rows);
}


#define DISPATCH_ROWS(rows, KERNEL, ...)                                              \
    switch (rows) {                                                                   \
    case 1: launch_k((KERNEL<1>), __VA_ARGS__); break;                      \
    case 2: launch_k((KERNEL<2>), __VA_ARGS__); break;                      \
    case 4: launch_k((KERNEL<4>), __VA_ARGS__); break;                      \
    case 8: launch_k((KERNEL<8>), __VA_ARGS__); break;                      \
    default: launch_k((KERNEL<16>), __VA_ARGS__); break;                    \
    }

// DISPATCH_ROWS for kernels templated on <rows, packed links>.
#define DISPATCH_ROWS_PK_(rows, KERNEL, PK, ...)                                     \
    switch (rows) {                                                                   \
    case 1: launch_k((KERNEL<1, PK>), __VA_ARGS__); break;                            \
    case 2: launch_k((KERNEL<2, PK>), __VA_ARGS__); break;                            \
    case 4: launch_k((KERNEL<4, PK>), __VA_ARGS__); break;                            \
    case 8: launch_k((KERNEL<8, PK>), __VA_ARGS__); break;                            \
    default: launch_k((KERNEL<16, PK>), __VA_ARGS__); break;                          \
    }
#define DISPATCH_ROWS_PK(rows, packed, KERNEL, ...)                                   \
    do {                                                                              \
        if (packed) {                                                                 \
            DISPATCH_ROWS_PK_(rows, KERNEL, true, __VA_ARGS__)                        \
        } else {                                                                      \
            DISPATCH_ROWS_PK_(rows, KERNEL, false, __VA_ARGS__)                       \
        }                                                                             \
    } while (0)

// Sums over the block, valid in thread 0 (a producer's per-block partial):
// one barrier, the waves' sums added in wave order (as block_allsum).  At most
// one call (of either) per kernel: they share their LDS.
__device__ __forceinline__ void block_sum2_t0(double a, double b, double &sa, double &sb)
{
    __shared__ double red[2 * (kBlock / kWave)];
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (lane == 0) {
        red[2 * w] = a;
        red[2 * w + 1] = b;
    }
    __syncthreads();
    sa = 0.0;
    sb = 0.0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / kWave;
        for (int i = 0; i < nw; ++i) {
            sa += red[2 * i];
            sb += red[2 * i + 1];
        }
    }
}
__device__ __forceinline__ double block_sum_t0(double v)
{
    double s, z;
    block_sum2_t0(v, 0.0, s, z);
    return s;
}


// Sum of the per-block partials part[b * stride + off], b < n, in a fixed
// order: lane l adds part[l], part[l + 64], ... in turn, then the lanes are
// summed by wave_sum.  Every wave of every block computes it by itself (no
// LDS, no barrier) and gets the same bits.  The partials were written by an
// earlier kernel, so the kernel boundary has made them visible.  (lds is not
// used.)
__device__ __forceinline__ double sum_prev_partials(const double *part, int n, int stride,
                                                    int off, double *)
{
    const int lane = threadIdx.x & (kWave - 1);
    double s = 0.0;
    int b = lane;
    for (; b + 3 * kWave < n; b += 4 * kWave) {
        const double v0 = part[(long)b * stride + off];
        const double v1 = part[(long)(b + kWave) * stride + off];
        const double v2 = part[(long)(b + 2 * kWave) * stride + off];
        const double v3 = part[(long)(b + 3 * kWave) * stride + off];
        s += v0;
        s += v1;
        s += v2;
        s += v3;
    }
    for (; b < n; b += kWave) s += part[(long)b * stride + off];
    return wave_sum(s);
}

// sum_prev_partials split in two: load() issues the loads (lane l holds
// part[l], part[l + 64], ...; lanes past n load part[0] and ignore it, so no
// load waits on a branch), sum() adds them in the same order (an ignored
// value adds +0.0, which leaves a sum of nonnegative partials unchanged) and
// sums the lanes: the same bits as sum_prev_partials.  A kernel issues the
// loads first and sums where the value is needed, with other loads in flight
// meanwhile.  For n <= kPartMax (a consumer never sees more: larger
// producers are folded).
constexpr int kPartMax = 1024;
constexpr int kPartLane = kPartMax / kWave;
struct PartialsAhead {
    double v[kPartLane];
    int n;
    __device__ __forceinline__ void load(const double *part, int n_, int stride, int off)
    {
        n = n_;
        const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
        for (int i = 0; i < kPartLane; ++i) {
            const int b = lane + i * kWave;
            v[i] = part[(long)(b < n ? b : 0) * stride + off];
        }
    }
    __device__ __forceinline__ double sum() const
    {
        const int lane = threadIdx.x & (kWave - 1);
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < kPartLane; ++i) s += (lane + i * kWave < n) ? v[i] : 0.0;
        return wave_sum(s);
    }
};
// Two interleaved partials per block (part[2b], part[2b + 1]), one 16-byte
// load each.
struct PartialsAhead2 {
    double2 v[kPartLane];
    int n;
    __device__ __forceinline__ void load(const double *part, int n_)
    {
        n = n_;
        const int lane = threadIdx.x & (kWave - 1);
        const double2 *p2 = reinterpret_cast<const double2 *>(part);
#pragma unroll
        for (int i = 0; i < kPartLane; ++i) {
            const int b = lane + i * kWave;
            v[i] = p2[b < n ? b : 0];
        }
    }
    __device__ __forceinline__ void sum(double &s0, double &s1) const
    {
        const int lane = threadIdx.x & (kWave - 1);
        double a = 0.0, c = 0.0;
#pragma unroll
        for (int i = 0; i < kPartLane; ++i) {
            const bool in = lane + i * kWave < n;
            a += in ? v[i].x : 0.0;
            c += in ? v[i].y : 0.0;
        }
        s0 = wave_sum(a);
        s1 = wave_sum(c);
    }
};

}  // namespace qed2d
