// qed2d_device.h -- device-side building blocks shared by the QED2d HIP kernels.
//
// Field layout in HBM (identical to the reference's Julia column-major
// [L1,L2,2] arrays, so caller buffers are used in place):
//   spinor / link   : T2[2][L2][L1]  (component or direction plane, then rows)
//   momentum / force: T [2][L2][L1]
//   site density    : T [L2][L1]
// with T2 = double2 (FP64) or float2 (FP32 shadows), (x, y) = (re, im).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qed2d {

constexpr int kWave = 64;

struct Latt {
    int L1, L2;
    long V;
    int bc;  // bit 0: fermion fields antiperiodic across i2 = L2-1 -> 0
};

// The FP32 kernels march site pairs (i1 = 2j, 2j+1) on lattices with even L1 >= 4.
inline bool pair_ok(const Latt &L) { return (L.L1 % 2) == 0 && L.L1 >= 4; }

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using c2 = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ c2<T> mk(T re, T im) { c2<T> z; z.x = re; z.y = im; return z; }
template <typename T> __device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ c2<T> cneg(c2<T> a) { return mk<T>(-a.x, -a.y); }
template <typename T> __device__ __forceinline__ c2<T> cconj(c2<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename T> __device__ __forceinline__ c2<T> cmulc(c2<T> a, c2<T> b)
{
    return mk<T>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename T> __device__ __forceinline__ c2<T> cscale(T s, c2<T> a) { return mk<T>(s * a.x, s * a.y); }
// i * a
template <typename T> __device__ __forceinline__ c2<T> cmuli(c2<T> a) { return mk<T>(-a.y, a.x); }
template <typename T> __device__ __forceinline__ T cabs2(c2<T> a) { return a.x * a.x + a.y * a.y; }

// Non-temporal-free plain loads; the compiler emits global_load_dwordx4 for
// double2 and global_load_dwordx2 for float2.
template <typename T> __device__ __forceinline__ c2<T> ld(const c2<T> *p, long i) { return p[i]; }

// ---------------------------------------------------------------------------
// The Wilson operator Q = gamma5 Dw at one site, from the values it touches
// (reference src/QED2dDirac.jl:2-33).  s*: (component 1, component 2) of the
// spinor at x (c), x+1 (p1), x-1 (m1), x+2 (p2), x-2 (m2).  U1 = U(x,1),
// U1m = U(x-1,1), U2 = U(x,2), U2m = U(x-2,2) (the last two carry the
// antiperiodic sign, if any).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void q_site(T mass2, c2<T> c1, c2<T> c2_, c2<T> p11, c2<T> p12,
                                       c2<T> m11, c2<T> m12, c2<T> p21, c2<T> p22,
                                       c2<T> m21, c2<T> m22, c2<T> U1, c2<T> U1m, c2<T> U2,
                                       c2<T> U2m, c2<T> &o1, c2<T> &o2)
{
    // forward/backward hops in direction 1
    c2<T> f1 = cmul<T>(U1, csub<T>(p11, p12));        // U1 (s1 - s2)(x+1)
    c2<T> b1 = cmulc<T>(U1m, cadd<T>(m11, m12));      // U1*(x-1) (s1 + s2)(x-1)
    // direction 2
    c2<T> f2a = cmul<T>(U2, cadd<T>(p21, cmuli<T>(p22)));   // U2 (s1 + i s2)(x+2)
    c2<T> b2a = cmulc<T>(U2m, csub<T>(m21, cmuli<T>(m22))); // U2*(x-2) (s1 - i s2)(x-2)
    c2<T> f2b = cmul<T>(U2, csub<T>(cmuli<T>(p21), p22));   // U2 (i s1 - s2)(x+2)
    c2<T> b2b = cmulc<T>(U2m, cadd<T>(m22, cmuli<T>(m21))); // U2*(x-2) (s2 + i s1)(x-2)
    const T h = T(0.5);
    c2<T> A = cscale<T>(h, cadd<T>(f1, b1));
    c2<T> B = cscale<T>(h, cadd<T>(f2a, b2a));
    c2<T> A2 = cscale<T>(h, csub<T>(f1, b1));
    c2<T> B2 = cscale<T>(h, csub<T>(f2b, b2b));
    o1 = csub<T>(csub<T>(cscale<T>(mass2, c1), A), B);
    o2 = csub<T>(csub<T>(cscale<T>(-mass2, c2_), A2), B2);
}

// ---------------------------------------------------------------------------
// Block-level reduction helpers (wave64 butterfly via shfl_xor, then LDS).
// The order of the tree is fixed by the launch shape, so results are
// reproducible run to run.
// ---------------------------------------------------------------------------
// One DPP data movement of a double (two 32-bit halves); lanes outside
// ROWMASK receive 0.
template <int CTRL, int ROWMASK = 0xF> __device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over the 64 lanes of a wave, returned to every lane.  Within each row of
// 16 lanes: quad swaps (xor 1, xor 2) and row rotations by 4 and 8 leave the
// row's sum in every lane; row_bcast:15 / row_bcast:31 then accumulate the
// rows into lane 63, which is broadcast.  A fixed order of additions: every
// wave that sums the same values gets the same bits.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);       // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);       // quad_perm [2,3,0,1]
    v += dpp_f64<0x124>(v);      // row_ror:4
    v += dpp_f64<0x128>(v);      // row_ror:8
    v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 -> rows 1, 3
    v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 -> rows 2, 3
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, kWave - 1);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), kWave - 1);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sums v over the block; result valid in thread 0.  lds must hold blockDim/64
// doubles.
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < nw; ++i) s += lds[i];
    }
    __syncthreads();
    return s;
}

// Double-double (hi, lo) accumulation (Knuth TwoSum) for action sums whose
// magnitude is ~V and whose differences must be resolved to ~1e-10.
struct dd { double hi, lo; };
__device__ __host__ __forceinline__ dd dd_add(dd a, double b)
{
    double s = a.hi + b;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b - bb);
    dd r;
    r.hi = s;
    r.lo = a.lo + e;
    return r;
}
__device__ __host__ __forceinline__ dd dd_add(dd a, dd b)
{
    dd r = dd_add(a, b.hi);
    r.lo += b.lo;
    double s = r.hi + r.lo;
    r.lo = r.lo - (s - r.hi);
    r.hi = s;
    return r;
}

__device__ __forceinline__ dd wave_sum_dd(dd v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        dd w;
        w.hi = __shfl_xor(v.hi, o, kWave);
        w.lo = __shfl_xor(v.lo, o, kWave);
        v = dd_add(v, w);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Counter-based random numbers: Philox4x32-10 (Salmon et al., SC'11) and
// Box-Muller with log / sin / cos evaluated by fixed FMA sequences.  The
// same sequences are restated in the CPU oracle, so host and device streams
// agree bit for bit.
// ---------------------------------------------------------------------------
__device__ __host__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    uint32_t c0 = c[0], c1 = c[1], c2v = c[2], c3 = c[3];
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2v;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2v = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    c[0] = c0; c[1] = c1; c[2] = c2v; c[3] = c3;
}

__device__ __host__ inline double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1.0p-53;
}

__device__ __host__ inline double det_log(double u)
{
#pragma clang fp contract(off)
    int e;
    double m = frexp(u, &e);
    if (m < 0x1.6a09e667f3bcdp-1) { m = m * 2.0; e -= 1; }
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    double p = 1.0 / 23.0;
    p = fma(p, s, 1.0 / 21.0);
    p = fma(p, s, 1.0 / 19.0);
    p = fma(p, s, 1.0 / 17.0);
    p = fma(p, s, 1.0 / 15.0);
    p = fma(p, s, 1.0 / 13.0);
    p = fma(p, s, 1.0 / 11.0);
    p = fma(p, s, 1.0 / 9.0);
    p = fma(p, s, 1.0 / 7.0);
    p = fma(p, s, 1.0 / 5.0);
    p = fma(p, s, 1.0 / 3.0);
    p = fma(p, s, 1.0);
    double t = 2.0 * (f * p);
    double de = (double)e;
    return fma(de, 0x1.62e42fee00000p-1, fma(de, 0x1.a39ef35793c76p-33, t));
}

__device__ __host__ inline void det_sincos2pi(double u, double *s_out, double *c_out)
{
#pragma clang fp contract(off)
    double y = 4.0 * u;
    double n = rint(y);
    double r = y - n;
    double z = r * r;
    double sp = -0x1.8a404211f9547p-45;
    sp = fma(sp, z, 0x1.aaec32af93359p-38);
    sp = fma(sp, z, -0x1.6fadb9f155744p-31);
    sp = fma(sp, z, 0x1.e8f434d018d63p-25);
    sp = fma(sp, z, -0x1.e3074fde8871fp-19);
    sp = fma(sp, z, 0x1.50783487ee782p-13);
    sp = fma(sp, z, -0x1.32d2cce62bd86p-8);
    sp = fma(sp, z, 0x1.466bc6775aae2p-4);
    sp = fma(sp, z, -0x1.4abbce625be53p-1);
    sp = fma(sp, z, 0x1.921fb54442d18p+0);
    double sv = sp * r;
    double cp = 0x1.ef6e308d6d1c4p-49;
    cp = fma(cp, z, -0x1.2a0c591af8314p-41);
    cp = fma(cp, z, 0x1.20c62c2f2d7f5p-34);
    cp = fma(cp, z, -0x1.b6e24f44b128fp-28);
    cp = fma(cp, z, 0x1.f9d38a3763cc3p-22);
    cp = fma(cp, z, -0x1.a6d1f2a204a8cp-16);
    cp = fma(cp, z, 0x1.e1f506891babbp-11);
    cp = fma(cp, z, -0x1.55d3c7e3cbffap-6);
    cp = fma(cp, z, 0x1.03c1f081b5ac4p-2);
    cp = fma(cp, z, -0x1.3bd3cc9be45dep+0);
    double cv = fma(cp, z, 1.0);
    int q = ((int)n) & 3;
    if (q == 0) { *s_out = sv; *c_out = cv; }
    else if (q == 1) { *s_out = cv; *c_out = -sv; }
    else if (q == 2) { *s_out = -sv; *c_out = -cv; }
    else { *s_out = -cv; *c_out = sv; }
}

// Gaussian pair from one Philox block: counter (idx, sub, purpose, traj),
// key (seed low word, seed high word ^ chain).
__device__ __host__ inline void gauss_pair(uint32_t idx, uint32_t sub, uint32_t purpose,
                                           uint32_t traj, uint32_t k0, uint32_t k1,
                                           double *g1, double *g2)
{
#pragma clang fp contract(off)
    uint32_t c[4] = {idx, sub, purpose, traj};
    philox4x32_10(c, k0, k1);
    double a = u53(c[0], c[1]);
    double b = u53(c[2], c[3]);
    double rad = sqrt(-2.0 * det_log(1.0 - a));
    double s, co;
    det_sincos2pi(b, &s, &co);
    *g1 = rad * co;
    *g2 = rad * s;
}

enum : uint32_t { kPurposeMom = 0x10, kPurposePF = 0x20, kPurposeMetro = 0x30, kPurposeHot = 0x40,
                  kPurposeVec = 0x50, kPurposeRew = 0x60 };

// XCD-aware block order: the hardware deals consecutive workgroups to the 8
// XCDs round robin; this mapping gives every XCD a contiguous band of tiles
// so that vertically adjacent tiles share one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int nx = 8;
    int per = (nb + nx - 1) / nx;
    int xcd = b % nx, k = b / nx;
    int t = xcd * per + k;
    // The last XCD band may be short; fall back to identity past the end.
    return (t < nb && (nb % nx == 0)) ? t : b;
}

}  // namespace qed2d
