// qed2d_kernels.hip -- HIP kernels of the QED2d molecular-dynamics path (gfx950).
//
// Every kernel here is HBM-bandwidth bound (the Wilson stencil does ~0.6
// flop/byte in FP64), so the design goals are: one pass over each field per
// logical use, coalesced 16-byte (FP64) or 8-byte (FP32) accesses along i1,
// vertical neighbours kept in registers by marching each thread up a column
// of `R` rows, and reductions finished on the device by the last block to
// arrive so that no kernel boundary needs a host round trip.
#include "qed2d_internal.h"

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
// Q apply: so = gamma5 Dw si  (src/QED2dDirac.jl:2-33)
// ===========================================================================
template <typename T> struct EpiStore {
    c2<T> *__restrict__ so;
    long V;
    __device__ __forceinline__ void operator()(long s, c2<T> o1, c2<T> o2, c2<T>, c2<T>)
    {
        so[s] = o1;
        so[V + s] = o2;
    }
};

template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_q_apply(Latt L, const c2<T> *__restrict__ U,
                                                    const c2<T> *__restrict__ si,
                                                    c2<T> *__restrict__ so, T mass2, int bx,
                                                    int by, int nbx)
{
    MarchPos p = march_pos(L, bx, by, nbx, R);
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
    __device__ __forceinline__ void operator()(long s, double2 o1, double2 o2, double2, double2)
    {
        double2 x1 = x[s], x2 = x[V + s];
        so[s] = mk<double>(o1.x + mu2 * x1.x, o1.y + mu2 * x1.y);
        so[V + s] = mk<double>(o2.x + mu2 * x2.x, o2.y + mu2 * x2.y);
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

#define DISPATCH_ROWS(rows, KERNEL, ...)                                              \
    switch (rows) {                                                                   \
    case 1: hipLaunchKernelGGL((KERNEL<1>), __VA_ARGS__); break;                      \
    case 2: hipLaunchKernelGGL((KERNEL<2>), __VA_ARGS__); break;                      \
    case 4: hipLaunchKernelGGL((KERNEL<4>), __VA_ARGS__); break;                      \
    case 8: hipLaunchKernelGGL((KERNEL<8>), __VA_ARGS__); break;                      \
    default: hipLaunchKernelGGL((KERNEL<16>), __VA_ARGS__); break;                    \
    }

template <typename T>
static void q_apply_impl(const Latt &L, const c2<T> *U, const c2<T> *si, c2<T> *so, T am0,
                         int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    const T mass2 = T(2) + am0;
    dim3 grid(g.nblocks()), blk(kBlock);
    switch (rows) {
    case 1: hipLaunchKernelGGL((k_q_apply<T, 1>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    case 2: hipLaunchKernelGGL((k_q_apply<T, 2>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    case 4: hipLaunchKernelGGL((k_q_apply<T, 4>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    case 8: hipLaunchKernelGGL((k_q_apply<T, 8>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
    default: hipLaunchKernelGGL((k_q_apply<T, 16>), grid, blk, 0, st, L, U, si, so, mass2, g.bx, g.by, g.nbx); break;
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
template <typename T>
__global__ void __launch_bounds__(kBlock) k_q_mrhs(Latt L, const c2<T> *__restrict__ U,
                                                   const c2<T> *__restrict__ si,
                                                   c2<T> *__restrict__ so, int nrhs, T mass2,
                                                   int bx, int by, int nbx)
{
    MarchPos p = march_pos(L, bx, by, nbx, 1);
    if (!p.active) return;
    const int L1 = L.L1, L2 = L.L2;
    const long V = L.V;
    const int i1 = p.i1, r = p.r0;
    const int i1p = (i1 + 1 == L1) ? 0 : i1 + 1;
    const int i1m = (i1 == 0) ? L1 - 1 : i1 - 1;
    const int rp = (r + 1 == L2) ? 0 : r + 1;
    const int rm = (r == 0) ? L2 - 1 : r - 1;
    const bool ap = (L.bc & 1) != 0;
    const long s = (long)r * L1 + i1;
    const c2<T> U1 = U[s], U1m = U[(long)r * L1 + i1m];
    c2<T> U2 = U[V + s], U2m = U[V + (long)rm * L1 + i1];
    if (ap && r == L2 - 1) U2 = cneg<T>(U2);
    if (ap && r == 0) U2m = cneg<T>(U2m);
    const long sp1 = (long)r * L1 + i1p, sm1 = (long)r * L1 + i1m;
    const long sp2 = (long)rp * L1 + i1, sm2 = (long)rm * L1 + i1;
    for (int j = 0; j < nrhs; ++j) {
        const c2<T> *in = si + (long)j * 2 * V;
        c2<T> *out = so + (long)j * 2 * V;
        c2<T> o1, o2;
        q_site<T>(mass2, in[s], in[V + s], in[sp1], in[V + sp1], in[sm1], in[V + sm1], in[sp2],
                  in[V + sp2], in[sm2], in[V + sm2], U1, U1m, U2, U2m, o1, o2);
        out[s] = o1;
        out[V + s] = o2;
    }
}

template <typename T>
static void q_mrhs_impl(const Latt &L, const c2<T> *U, const c2<T> *si, c2<T> *so, int nrhs,
                        T am0, hipStream_t st)
{
    MarchGrid g = march_grid(L, 1);
    hipLaunchKernelGGL(k_q_mrhs<T>, dim3(g.nblocks()), dim3(kBlock), 0, st, L, U, si, so, nrhs,
                       T(2) + am0, g.bx, g.by, g.nbx);
}
void launch_q_apply_mrhs_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                             int nrhs, double am0, hipStream_t st)
{
    q_mrhs_impl<double>(L, U, si, so, nrhs, am0, st);
}
void launch_q_apply_mrhs_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so,
                             int nrhs, float am0, hipStream_t st)
{
    q_mrhs_impl<float>(L, U, si, so, nrhs, am0, st);
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
    hipLaunchKernelGGL(k_gamm5, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, si, so, L.V);
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
    hipLaunchKernelGGL(k_axpy_c, dim3(ew_blocks(n)), dim3(kBlock), 0, st, y, x, ar, ai, n);
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
    hipLaunchKernelGGL(k_scale_add, dim3(ew_blocks(n)), dim3(kBlock), 0, st, y, a, x, b, n);
}

__global__ void k_copy(double2 *__restrict__ d, const double2 *__restrict__ s, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        d[i] = s[i];
}
void launch_copy(double2 *dst, const double2 *src, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy, dim3(ew_blocks(n)), dim3(kBlock), 0, st, dst, src, n);
}

__global__ void k_negate(double *a, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        a[i] = -a[i];
}
void launch_negate(double *a, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_negate, dim3(ew_blocks(n)), dim3(kBlock), 0, st, a, n);
}

__global__ void k_to_f32(const double2 *__restrict__ in, float2 *__restrict__ out, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 v = in[i];
        out[i] = mk<float>((float)v.x, (float)v.y);
    }
}
void launch_to_f32(const double2 *in, float2 *out, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f32, dim3(ew_blocks(n)), dim3(kBlock), 0, st, in, out, n);
}

// ===========================================================================
// Deterministic reductions: fixed grid, fixed tree, last block finishes.
// ===========================================================================
constexpr int kRedBlocks = 1024;

static inline int red_blocks(long n)
{
    long b = (n + kBlock - 1) / kBlock;
    if (b > kRedBlocks) b = kRedBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

// a^dag b (complex) or |a|^2 when b == nullptr.
__global__ void __launch_bounds__(kBlock) k_dot(const double2 *__restrict__ a,
                                                const double2 *__restrict__ b, long n,
                                                RedOut *out, double *part)
{
    __shared__ double lds[kBlock / kWave];
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
    double br = block_sum(sr, lds);
    double bi = block_sum(si, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = br;
        part[2 * blockIdx.x + 1] = bi;
    }
    if (last_block_arrives<2>(&out->ticket[0], gridDim.x)) {
        double res[2];
        sum_partials<2>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            out->v[0] = res[0];
            out->v[1] = res[1];
            out->ticket[0] = 0u;
        }
    }
}
void launch_norm2(const double2 *a, long n, RedOut *out, double *part, hipStream_t st)
{
    hipLaunchKernelGGL(k_dot, dim3(red_blocks(n)), dim3(kBlock), 0, st, a, nullptr, n, out, part);
}
void launch_dot(const double2 *a, const double2 *b, long n, RedOut *out, double *part,
                hipStream_t st)
{
    hipLaunchKernelGGL(k_dot, dim3(red_blocks(n)), dim3(kBlock), 0, st, a, b, n, out, part);
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

// Plaquette P(x) = U1(x) U2(x+1) conj(U1(x+2) U2(x))  (src/QED2dAction.jl:10-13)
__device__ __forceinline__ double2 plaq_at(const double2 *__restrict__ U, const Latt &L, int i1,
                                           int i2)
{
    const int iu1 = (i1 + 1 == L.L1) ? 0 : i1 + 1;
    const int iu2 = (i2 + 1 == L.L2) ? 0 : i2 + 1;
    const long V = L.V;
    double2 a = cmul<double>(U[(long)i2 * L.L1 + i1], U[V + (long)i2 * L.L1 + iu1]);
    double2 b = cmul<double>(U[(long)iu2 * L.L1 + i1], U[V + (long)i2 * L.L1 + i1]);
    return cmul<double>(a, cconj<double>(b));
}

// Sum Re P, sum arg P (src/QED2dAction.jl:2-32) and sum p^2 (src/QED2dHMC.jl:200),
// each as a double-double; optional per-site density outputs.
__global__ void __launch_bounds__(kBlock) k_observables(Latt L, const double2 *__restrict__ U,
                                                        const double *__restrict__ mom,
                                                        double *__restrict__ plx,
                                                        double *__restrict__ qdens,
                                                        RedOut *out, double *part)
{
    __shared__ double lds[2 * (kBlock / kWave)];
    dd sp = {0.0, 0.0}, sq = {0.0, 0.0}, sm = {0.0, 0.0};
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)gridDim.x * blockDim.x) {
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
        double *pp = part + 6 * (long)blockIdx.x;
        pp[0] = bp.hi; pp[1] = bp.lo; pp[2] = bq.hi; pp[3] = bq.lo; pp[4] = bm.hi; pp[5] = bm.lo;
    }
    if (last_block_arrives<6>(&out->ticket[0], gridDim.x)) {
        for (int k = 0; k < 3; ++k) {
            dd acc = {0.0, 0.0};
            for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
                dd t = {__builtin_nontemporal_load(&part[6 * (long)b + 2 * k]),
                        __builtin_nontemporal_load(&part[6 * (long)b + 2 * k + 1])};
                acc = dd_add(acc, t);
            }
            acc = block_sum_dd(acc, lds);
            if (threadIdx.x == 0) {
                out->v[2 * k] = acc.hi;
                out->v[2 * k + 1] = acc.lo;
            }
        }
        if (threadIdx.x == 0) out->ticket[0] = 0u;
    }
}
void launch_observables(const Latt &L, const double2 *U, const double *mom, double *plx,
                        double *qdens, RedOut *out, double *part, hipStream_t st)
{
    hipLaunchKernelGGL(k_observables, dim3(red_blocks(L.V)), dim3(kBlock), 0, st, L, U, mom, plx,
                       qdens, out, part);
}

// ===========================================================================
// Conjugate gradient on Q^2 + mu^2 (src/QED2dSolver.jl:3-105), restructured
// into two fused kernels per iteration:
//   K1: p = r + beta p_old (on the fly), t = Q p, |t|^2, |p|^2
//   K2: Ap = Q t + mu^2 p, alpha = |r|^2 / (|t|^2 + mu^2 |p|^2),
//       x += alpha p, r -= alpha Ap, |r|^2, convergence test
// p^dag (Q^2 + mu^2) p = |Q p|^2 + mu^2 |p|^2 because Q is Hermitian, so the
// reference's dot(p, Ap) needs no separate pass.  The scalars live in
// SolverState; converged solves turn the remaining queued kernels into no-ops.
// ===========================================================================
template <typename T> struct EpiK1 {
    c2<T> *__restrict__ pnew;
    c2<T> *__restrict__ t;
    long V;
    double tt, pp;
    __device__ __forceinline__ void operator()(long s, c2<T> o1, c2<T> o2, c2<T> c1, c2<T> cc2)
    {
        pnew[s] = c1;
        pnew[V + s] = cc2;
        t[s] = o1;
        t[V + s] = o2;
        tt += (double)(o1.x * o1.x + o1.y * o1.y) + (double)(o2.x * o2.x + o2.y * o2.y);
        pp += (double)(c1.x * c1.x + c1.y * c1.y) + (double)(cc2.x * cc2.x + cc2.y * cc2.y);
    }
};

// The two search-direction buffers alternate with the device's iteration
// count: iteration k = iter + 1 reads P[iter & 1] and writes P[(iter + 1) & 1].
// Selecting them on the device keeps the pairing right when queued
// iterations are skipped (converged, or waiting for a reliable update).
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_cg1(Latt L, const c2<T> *__restrict__ U,
                                                const c2<T> *__restrict__ r, c2<T> *P0,
                                                c2<T> *P1, c2<T> *__restrict__ t, T mass2,
                                                SolverState *S, double *part, int bx, int by,
                                                int nbx)
{
    __shared__ double lds[kBlock / kWave];
    if (S->status != 0 || S->need_ru != 0) return;
    const bool odd = (S->iter & 1) != 0;
    const c2<T> *pold = odd ? P1 : P0;
    c2<T> *pnew = odd ? P0 : P1;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    EpiK1<T> epi{pnew, t, L.V, 0.0, 0.0};
    if (p.active) {
        SrcXpay<T> src{r, pold, (T)S->beta, L.V};
        q_march<T, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a = block_sum(epi.tt, lds);
    double b = block_sum(epi.pp, lds);
    if (threadIdx.x == 0) {
        part[2 * p.blk] = a;
        part[2 * p.blk + 1] = b;
    }
    if (last_block_arrives<2>(&S->ticket[0], gridDim.x)) {
        double res[2];
        sum_partials<2>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            S->tt = res[0];
            S->pp = res[1];
            S->ticket[0] = 0u;
        }
    }
}

template <typename T> struct EpiK2 {
    const c2<T> *__restrict__ p;
    c2<T> *__restrict__ x;
    c2<T> *__restrict__ r;
    T alpha, mu2;
    long V;
    double rr;
    __device__ __forceinline__ void operator()(long s, c2<T> o1, c2<T> o2, c2<T>, c2<T>)
    {
        c2<T> p1 = p[s], p2 = p[V + s];
        c2<T> a1 = mk<T>(o1.x + mu2 * p1.x, o1.y + mu2 * p1.y);
        c2<T> a2 = mk<T>(o2.x + mu2 * p2.x, o2.y + mu2 * p2.y);
        c2<T> x1 = x[s], x2 = x[V + s], r1 = r[s], r2 = r[V + s];
        x1 = mk<T>(x1.x + alpha * p1.x, x1.y + alpha * p1.y);
        x2 = mk<T>(x2.x + alpha * p2.x, x2.y + alpha * p2.y);
        r1 = mk<T>(r1.x - alpha * a1.x, r1.y - alpha * a1.y);
        r2 = mk<T>(r2.x - alpha * a2.x, r2.y - alpha * a2.y);
        x[s] = x1;
        x[V + s] = x2;
        r[s] = r1;
        r[V + s] = r2;
        rr += (double)(r1.x * r1.x + r1.y * r1.y) + (double)(r2.x * r2.x + r2.y * r2.y);
    }
};

template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_cg2(Latt L, const c2<T> *__restrict__ U,
                                                const c2<T> *__restrict__ t, const c2<T> *P0,
                                                const c2<T> *P1, c2<T> *__restrict__ x,
                                                c2<T> *__restrict__ r, T mass2, SolverState *S,
                                                double *part, int bx, int by, int nbx, int mixed)
{
    __shared__ double lds[kBlock / kWave];
    if (S->status != 0 || S->need_ru != 0) return;
    const c2<T> *pv = (S->iter & 1) ? P0 : P1;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    const double alpha = S->norm / (S->tt + S->mu2 * S->pp);
    EpiK2<T> epi{pv, x, r, (T)alpha, (T)S->mu2, L.V, 0.0};
    if (p.active) {
        SrcPlain<T> src{t, L.V};
        q_march<T, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a = block_sum(epi.rr, lds);
    if (threadIdx.x == 0) part[p.blk] = a;
    if (last_block_arrives<1>(&S->ticket[1], gridDim.x)) {
        double res[1];
        sum_partials<1>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            const double err = res[0];
            S->err = err;
            S->iter += 1;
            if (!mixed) {
                // src/QED2dSolver.jl:27-37
                if (err < S->tol) {
                    S->status = 1;
                } else {
                    S->beta = err / S->norm;
                    S->norm = err;
                    if (S->iter >= S->maxiter) S->status = 2;
                }
            } else {
                // FP32 iteration: ask for a reliable update once the iterated
                // residual has dropped by delta since the last one, or looks
                // converged.
                if (err < S->tol || err < S->delta2 * S->rmax2 || S->iter >= S->maxiter) {
                    S->need_ru = 1;
                } else {
                    S->beta = err / S->norm;
                    S->norm = err;
                    if (err > S->rmax2) S->rmax2 = err;
                }
            }
            S->ticket[1] = 0u;
        }
    }
}

template <typename T>
static void cg1_impl(const Latt &L, const c2<T> *U, const c2<T> *r, c2<T> *pold, c2<T> *pnew,
                     c2<T> *t, T am0, SolverState *S, double *part, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    const T m2 = T(2) + am0;
    switch (rows) {
    case 1: hipLaunchKernelGGL((k_cg1<T, 1>), grid, blk, 0, st, L, U, r, pold, pnew, t, m2, S, part, g.bx, g.by, g.nbx); break;
    case 2: hipLaunchKernelGGL((k_cg1<T, 2>), grid, blk, 0, st, L, U, r, pold, pnew, t, m2, S, part, g.bx, g.by, g.nbx); break;
    case 4: hipLaunchKernelGGL((k_cg1<T, 4>), grid, blk, 0, st, L, U, r, pold, pnew, t, m2, S, part, g.bx, g.by, g.nbx); break;
    case 8: hipLaunchKernelGGL((k_cg1<T, 8>), grid, blk, 0, st, L, U, r, pold, pnew, t, m2, S, part, g.bx, g.by, g.nbx); break;
    default: hipLaunchKernelGGL((k_cg1<T, 16>), grid, blk, 0, st, L, U, r, pold, pnew, t, m2, S, part, g.bx, g.by, g.nbx); break;
    }
}

template <typename T>
static void cg2_impl(const Latt &L, const c2<T> *U, const c2<T> *t, const c2<T> *p0,
                     const c2<T> *p1, c2<T> *x, c2<T> *r, T am0, SolverState *S, double *part,
                     int rows, int mixed, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    const T m2 = T(2) + am0;
    switch (rows) {
    case 1: hipLaunchKernelGGL((k_cg2<T, 1>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, part, g.bx, g.by, g.nbx, mixed); break;
    case 2: hipLaunchKernelGGL((k_cg2<T, 2>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, part, g.bx, g.by, g.nbx, mixed); break;
    case 4: hipLaunchKernelGGL((k_cg2<T, 4>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, part, g.bx, g.by, g.nbx, mixed); break;
    case 8: hipLaunchKernelGGL((k_cg2<T, 8>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, part, g.bx, g.by, g.nbx, mixed); break;
    default: hipLaunchKernelGGL((k_cg2<T, 16>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, part, g.bx, g.by, g.nbx, mixed); break;
    }
}

void launch_cg_k1_f64(const Latt &L, const double2 *U, const double2 *r, double2 *p0,
                      double2 *p1, double2 *t, double am0, SolverState *s, double *part,
                      int rows, hipStream_t st)
{
    cg1_impl<double>(L, U, r, p0, p1, t, am0, s, part, rows, st);
}
void launch_cg_k2_f64(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                      const double2 *p1, double2 *x, double2 *r, double am0, SolverState *s,
                      double *part, int rows, int mixed, hipStream_t st)
{
    cg2_impl<double>(L, U, t, p0, p1, x, r, am0, s, part, rows, mixed, st);
}
void launch_cg_k1_f32(const Latt &L, const float2 *U, const float2 *r, float2 *p0, float2 *p1,
                      float2 *t, float am0, SolverState *s, double *part, int rows,
                      hipStream_t st)
{
    cg1_impl<float>(L, U, r, p0, p1, t, am0, s, part, rows, st);
}
void launch_cg_k2_f32(const Latt &L, const float2 *U, const float2 *t, const float2 *p0,
                      const float2 *p1, float2 *x, float2 *r, float am0, SolverState *s,
                      double *part, int rows, int mixed, hipStream_t st)
{
    cg2_impl<float>(L, U, t, p0, p1, x, r, am0, s, part, rows, mixed, st);
}

// x = 0, r = b, p0 = 0, |b|^2 -> norm, bnorm, tol  (src/QED2dSolver.jl:10-16)
__global__ void __launch_bounds__(kBlock) k_cg_init(const double2 *__restrict__ b,
                                                    double2 *__restrict__ x,
                                                    double2 *__restrict__ r,
                                                    double2 *__restrict__ p0, long n,
                                                    SolverState *S, double *part)
{
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    const double2 z = mk<double>(0.0, 0.0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 v = b[i];
        x[i] = z;
        r[i] = v;
        p0[i] = z;
        acc += v.x * v.x + v.y * v.y;
    }
    double a = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
    if (last_block_arrives<1>(&S->ticket[2], gridDim.x)) {
        double res[1];
        sum_partials<1>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            S->bnorm = res[0];
            S->norm = res[0];
            S->tol = S->eps * res[0];
            S->rmax2 = res[0];
            S->beta = 0.0;
            S->iter = 0;
            S->need_ru = 0;
            S->n_ru = 0;
            // b = 0 has the exact solution x = 0.
            S->status = (res[0] == 0.0) ? 1 : 0;
            S->ticket[2] = 0u;
        }
    }
}
void launch_cg_init_f64(const Latt &L, const double2 *b, double2 *x, double2 *r, double2 *p0,
                        SolverState *s, double *part, hipStream_t st)
{
    const long n = 2 * L.V;
    hipLaunchKernelGGL(k_cg_init, dim3(red_blocks(n)), dim3(kBlock), 0, st, b, x, r, p0, n, s,
                       part);
}

// Residual of a starting guess: r = b - (Q t + mu^2 x), t = Q x
// (src/QED2dSolver.jl:65-72).  mode 0: FP64 r; mode 1 (mixed, initial) and
// mode 2 (mixed, reliable update) write the FP32 residual instead.
struct EpiResid {
    const double2 *__restrict__ b;
    const double2 *__restrict__ x;
    double2 *__restrict__ r;
    float2 *__restrict__ rs;
    double2 *__restrict__ p0;
    float2 *__restrict__ p0s;
    double mu2;
    long V;
    double rr, bb;
    __device__ __forceinline__ void one(long i, double2 o)
    {
        double2 bv = b[i], xv = x[i];
        double2 rv = mk<double>(bv.x - (o.x + mu2 * xv.x), bv.y - (o.y + mu2 * xv.y));
        if (r) r[i] = rv;
        if (rs) rs[i] = mk<float>((float)rv.x, (float)rv.y);
        if (p0) p0[i] = mk<double>(0.0, 0.0);
        if (p0s) p0s[i] = mk<float>(0.0f, 0.0f);
        rr += rv.x * rv.x + rv.y * rv.y;
        bb += bv.x * bv.x + bv.y * bv.y;
    }
    __device__ __forceinline__ void operator()(long s, double2 o1, double2 o2, double2, double2)
    {
        one(s, o1);
        one(V + s, o2);
    }
};

template <int R>
__global__ void __launch_bounds__(kBlock) k_resid(Latt L, const double2 *__restrict__ U,
                                                  const double2 *__restrict__ t,
                                                  EpiResid epi0, double mass2, SolverState *S,
                                                  double *part, int bx, int by, int nbx,
                                                  int mode)
{
    __shared__ double lds[kBlock / kWave];
    if (mode == 2 && (S->status != 0 || S->need_ru == 0)) return;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    EpiResid epi = epi0;
    epi.mu2 = S->mu2;
    epi.rr = 0.0;
    epi.bb = 0.0;
    if (p.active) {
        SrcPlain<double> src{t, L.V};
        q_march<double, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a = block_sum(epi.rr, lds);
    double b = block_sum(epi.bb, lds);
    if (threadIdx.x == 0) {
        part[2 * p.blk] = a;
        part[2 * p.blk + 1] = b;
    }
    if (last_block_arrives<2>(&S->ticket[3], gridDim.x)) {
        double res[2];
        sum_partials<2>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            const double r2 = res[0];
            if (mode != 2) {
                S->bnorm = res[1];
                S->tol = S->eps * res[1];
                S->norm = r2;
                S->rmax2 = r2;
                S->beta = 0.0;
                S->iter = 0;
                S->need_ru = 0;
                S->n_ru = 0;
                S->status = (r2 < S->tol || res[1] == 0.0) ? 1 : 0;
            } else {
                S->n_ru += 1;
                S->need_ru = 0;
                if (r2 < S->tol) {
                    S->status = 1;
                } else {
                    S->beta = r2 / S->norm;
                    S->norm = r2;
                    S->rmax2 = r2;
                    if (S->iter >= S->maxiter) S->status = 2;
                }
            }
            S->err = r2;
            S->ticket[3] = 0u;
        }
    }
}

static void resid_impl(const Latt &L, const double2 *U, const double2 *t, EpiResid e,
                       double am0, SolverState *S, double *part, int rows, int mode,
                       hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    DISPATCH_ROWS(rows, k_resid, grid, blk, 0, st, L, U, t, e, 2.0 + am0, S, part, g.bx, g.by,
                  g.nbx, mode);
}

void launch_cg_guess_f64(const Latt &L, const double2 *U, const double2 *b, const double2 *x,
                         const double2 *t, double2 *r, double2 *p0, double am0, SolverState *s,
                         double *part, int rows, hipStream_t st)
{
    EpiResid e{b, x, r, nullptr, p0, nullptr, 0.0, L.V, 0.0, 0.0};
    resid_impl(L, U, t, e, am0, s, part, rows, 0, st);
}

void launch_ru_residual(const Latt &L, const double2 *U, const double2 *b, const double2 *t,
                        const double2 *x, float2 *rs, float2 *p0, double am0, SolverState *s,
                        double *part, int rows, int initial, hipStream_t st)
{
    EpiResid e{b, x, nullptr, rs, nullptr, initial ? p0 : nullptr, 0.0, L.V, 0.0, 0.0};
    resid_impl(L, U, t, e, am0, s, part, rows, initial ? 1 : 2, st);
}

// Mixed-precision start from x = 0: rs = float(b), xs = 0, p0 = 0.
__global__ void __launch_bounds__(kBlock) k_cg_init_mixed(const double2 *__restrict__ b,
                                                          float2 *__restrict__ rs,
                                                          float2 *__restrict__ xs,
                                                          float2 *__restrict__ p0, long n,
                                                          SolverState *S, double *part)
{
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    const float2 z = mk<float>(0.0f, 0.0f);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double2 v = b[i];
        rs[i] = mk<float>((float)v.x, (float)v.y);
        xs[i] = z;
        p0[i] = z;
        acc += v.x * v.x + v.y * v.y;
    }
    double a = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
    if (last_block_arrives<1>(&S->ticket[2], gridDim.x)) {
        double res[1];
        sum_partials<1>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            S->bnorm = res[0];
            S->norm = res[0];
            S->tol = S->eps * res[0];
            S->rmax2 = res[0];
            S->beta = 0.0;
            S->iter = 0;
            S->need_ru = 0;
            S->n_ru = 0;
            S->status = (res[0] == 0.0) ? 1 : 0;
            S->ticket[2] = 0u;
        }
    }
}
void launch_cg_init_mixed(const Latt &L, const double2 *b, float2 *rs, float2 *xs, float2 *p0,
                          SolverState *s, double *part, hipStream_t st)
{
    const long n = 2 * L.V;
    hipLaunchKernelGGL(k_cg_init_mixed, dim3(red_blocks(n)), dim3(kBlock), 0, st, b, rs, xs, p0,
                       n, s, part);
}

// x += double(xs); xs = 0  (a reliable update folds the FP32 partial solution
// into the FP64 one).
__global__ void k_ru_accum(double2 *__restrict__ x, float2 *__restrict__ xs, long n,
                           const SolverState *S)
{
    if (S->need_ru == 0 || S->status != 0) return;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        float2 d = xs[i];
        double2 v = x[i];
        x[i] = mk<double>(v.x + (double)d.x, v.y + (double)d.y);
        xs[i] = mk<float>(0.0f, 0.0f);
    }
}
void launch_ru_accum(const Latt &L, double2 *x, float2 *xs, const SolverState *s,
                     hipStream_t st)
{
    hipLaunchKernelGGL(k_ru_accum, dim3(ew_blocks(2 * L.V)), dim3(kBlock), 0, st, x, xs, 2 * L.V,
                       s);
}

// ===========================================================================
// Multi-shift CG (Jegerlehner, hep-lat/9612014): the systems
// (Q^2 + sigma_i) x_i = b share the Krylov space of the base system
// (Q^2 + sigma_b), whose residual r_k fixes every shifted residual up to a
// scalar, r^i_k = zeta^i_k r_k.  Per iteration k:
//   K1: p_k = r_k + beta_k p_{k-1}, t = Q p_k, |t|^2, |p_k|^2, and for each
//       shift x_i += alpha^i_{k-1} p^i_{k-1}, p^i_k = zeta^i_k r_k + beta^i_k p^i_{k-1}
//   K2: Ap = Q t + sigma_b p_k, alpha_k = |r_k|^2 / p_k^dag Ap, r_{k+1} = r_k - alpha_k Ap,
//       |r_{k+1}|^2, then (last block) the zeta / alpha / beta recurrences:
//   zeta_{k+1} = zeta_k zeta_{k-1} alpha_{k-1} /
//                (alpha_k beta_k (zeta_{k-1} - zeta_k) + zeta_{k-1} alpha_{k-1} (1 + s alpha_k))
//   alpha^i_k = alpha_k zeta_{k+1} / zeta_k,   beta^i_{k+1} = beta_{k+1} (zeta_{k+1} / zeta_k)^2
// with s = sigma_i - sigma_b.  A shift stops once zeta^2 |r|^2 < tol.
// Algorithmic bytes per site and iteration (FP64): 320 + 128 n_active.
// ===========================================================================
struct EpiMS1 {
    c2<double> *__restrict__ pnew;
    c2<double> *__restrict__ t;
    const c2<double> *__restrict__ r;
    const MultiShiftState *S;
    int nshift;
    long V;
    double tt, pp;
    __device__ __forceinline__ void one(long i, double2 rv)
    {
        for (int k = 0; k < nshift; ++k) {
            const int a = S->active[k];
            if (a == 0) continue;
            double2 *ps = S->ps[k];
            double2 *x = S->x[k];
            const double2 pv = ps[i];
            const double as = S->alpha_s[k];
            double2 xv = x[i];
            x[i] = mk<double>(xv.x + as * pv.x, xv.y + as * pv.y);
            if (a == 1) {
                const double z = S->zeta[k], bs = S->beta_s[k];
                ps[i] = mk<double>(z * rv.x + bs * pv.x, z * rv.y + bs * pv.y);
            }
        }
    }
    __device__ __forceinline__ void operator()(long s, double2 o1, double2 o2, double2 c1,
                                               double2 cc2)
    {
        pnew[s] = c1;
        pnew[V + s] = cc2;
        t[s] = o1;
        t[V + s] = o2;
        tt += (o1.x * o1.x + o1.y * o1.y) + (o2.x * o2.x + o2.y * o2.y);
        pp += (c1.x * c1.x + c1.y * c1.y) + (cc2.x * cc2.x + cc2.y * cc2.y);
        one(s, r[s]);
        one(V + s, r[V + s]);
    }
};

template <int R>
__global__ void __launch_bounds__(kBlock) k_ms1(Latt L, const double2 *__restrict__ U,
                                                const double2 *__restrict__ r, double2 *P0,
                                                double2 *P1, double2 *__restrict__ t,
                                                double mass2, MultiShiftState *S, double *part,
                                                int nshift, int bx, int by, int nbx)
{
    __shared__ double lds[kBlock / kWave];
    if (S->status != 0) return;
    const bool odd = (S->iter & 1) != 0;
    const double2 *pold = odd ? P1 : P0;
    double2 *pnew = odd ? P0 : P1;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    EpiMS1 epi{pnew, t, r, S, nshift, L.V, 0.0, 0.0};
    if (p.active) {
        SrcXpay<double> src{r, pold, S->beta, L.V};
        q_march<double, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a = block_sum(epi.tt, lds);
    double b = block_sum(epi.pp, lds);
    if (threadIdx.x == 0) {
        part[2 * p.blk] = a;
        part[2 * p.blk + 1] = b;
    }
    if (last_block_arrives<2>(&S->ticket[0], gridDim.x)) {
        double res[2];
        sum_partials<2>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            S->tt = res[0];
            S->pp = res[1];
            for (int k = 0; k < nshift; ++k) {
                if (S->active[k] == 2) S->active[k] = 0;  // last x update done
                S->alpha_s[k] = 0.0;
            }
            S->ticket[0] = 0u;
        }
    }
}

struct EpiMS2 {
    double2 *__restrict__ r;
    const double2 *__restrict__ p;
    double alpha, sigma;
    long V;
    double rr;
    __device__ __forceinline__ void operator()(long s, double2 o1, double2 o2, double2, double2)
    {
        const double2 p1 = p[s], p2 = p[V + s];
        double2 r1 = r[s], r2 = r[V + s];
        r1 = mk<double>(r1.x - alpha * (o1.x + sigma * p1.x), r1.y - alpha * (o1.y + sigma * p1.y));
        r2 = mk<double>(r2.x - alpha * (o2.x + sigma * p2.x), r2.y - alpha * (o2.y + sigma * p2.y));
        r[s] = r1;
        r[V + s] = r2;
        rr += (r1.x * r1.x + r1.y * r1.y) + (r2.x * r2.x + r2.y * r2.y);
    }
};

template <int R>
__global__ void __launch_bounds__(kBlock) k_ms2(Latt L, const double2 *__restrict__ U,
                                                const double2 *__restrict__ t, const double2 *P0,
                                                const double2 *P1, double2 *__restrict__ r,
                                                double mass2, MultiShiftState *S, double *part,
                                                int nshift, int bx, int by, int nbx)
{
    __shared__ double lds[kBlock / kWave];
    if (S->status != 0) return;
    const double2 *pv = (S->iter & 1) ? P0 : P1;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    const double alpha = S->rr / (S->tt + S->sigma_b * S->pp);
    EpiMS2 epi{r, pv, alpha, S->sigma_b, L.V, 0.0};
    if (p.active) {
        SrcPlain<double> src{t, L.V};
        q_march<double, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a = block_sum(epi.rr, lds);
    if (threadIdx.x == 0) part[p.blk] = a;
    if (last_block_arrives<1>(&S->ticket[1], gridDim.x)) {
        double res[1];
        sum_partials<1>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            const double rr_new = res[0];
            const double beta_k = S->beta, alpha_p = S->alpha_prev;
            const double beta_n = rr_new / S->rr;
            int running = 0;
            for (int k = 0; k < nshift; ++k) {
                if (S->active[k] != 1) continue;
                const double z = S->zeta[k], zp = S->zeta_prev[k];
                const double zn = z * zp * alpha_p /
                                  (alpha * beta_k * (zp - z) + zp * alpha_p * (1.0 + S->s[k] * alpha));
                S->alpha_s[k] = alpha * zn / z;
                S->beta_s[k] = beta_n * (zn / z) * (zn / z);
                S->zeta_prev[k] = z;
                S->zeta[k] = zn;
                if (zn * zn * rr_new < S->tol) {
                    S->active[k] = 2;
                    S->conv_iter[k] = S->iter + 1;
                } else {
                    running = 1;
                }
            }
            S->alpha_prev = alpha;
            S->beta = beta_n;
            S->rr = rr_new;
            S->iter += 1;
            if (!running) S->status = 1;
            else if (S->iter >= S->maxiter) S->status = 2;
            S->ticket[1] = 0u;
        }
    }
}

// Pending x updates of the shifts that converged in the last iteration.
__global__ void k_ms_final(MultiShiftState *S, int nshift, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        for (int k = 0; k < nshift; ++k) {
            if (S->active[k] != 2) continue;
            const double as = S->alpha_s[k];
            const double2 pv = S->ps[k][i];
            double2 xv = S->x[k][i];
            S->x[k][i] = mk<double>(xv.x + as * pv.x, xv.y + as * pv.y);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_ms_init(const double2 *__restrict__ b,
                                                    double2 *__restrict__ r,
                                                    double2 *__restrict__ p0, long n,
                                                    MultiShiftState *S, double *part,
                                                    int nshift)
{
    __shared__ double lds[kBlock / kWave];
    double acc = 0.0;
    const double2 z = mk<double>(0.0, 0.0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double2 v = b[i];
        r[i] = v;
        p0[i] = z;
        for (int k = 0; k < nshift; ++k) {
            S->x[k][i] = z;
            S->ps[k][i] = z;
        }
        acc += v.x * v.x + v.y * v.y;
    }
    double a = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
    if (last_block_arrives<1>(&S->ticket[2], gridDim.x)) {
        double res[1];
        sum_partials<1>(part, gridDim.x, res, lds);
        if (threadIdx.x == 0) {
            S->bnorm = res[0];
            S->rr = res[0];
            S->tol = S->eps * res[0];
            S->alpha_prev = 1.0;
            S->beta = 0.0;
            S->iter = 0;
            for (int k = 0; k < nshift; ++k) {
                S->zeta[k] = 1.0;
                S->zeta_prev[k] = 1.0;
                S->alpha_s[k] = 0.0;
                S->beta_s[k] = 0.0;
                S->active[k] = 1;
                S->conv_iter[k] = 0;
            }
            S->status = (res[0] == 0.0) ? 1 : 0;
            if (res[0] == 0.0)
                for (int k = 0; k < nshift; ++k) S->active[k] = 0;
            S->ticket[2] = 0u;
        }
    }
}

void launch_ms_init(const Latt &L, const double2 *b, double2 *r, double2 *p0,
                    MultiShiftState *s, double *part, int nshift, hipStream_t st)
{
    const long n = 2 * L.V;
    hipLaunchKernelGGL(k_ms_init, dim3(red_blocks(n)), dim3(kBlock), 0, st, b, r, p0, n, s, part,
                       nshift);
}

void launch_ms_k1(const Latt &L, const double2 *U, const double2 *r, double2 *p0, double2 *p1,
                  double2 *t, double am0, MultiShiftState *s, double *part, int nshift, int rows,
                  hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    DISPATCH_ROWS(rows, k_ms1, grid, blk, 0, st, L, U, r, p0, p1, t, 2.0 + am0, s, part, nshift,
                  g.bx, g.by, g.nbx);
}

void launch_ms_k2(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                  const double2 *p1, double2 *r, double am0, MultiShiftState *s, double *part,
                  int nshift, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    DISPATCH_ROWS(rows, k_ms2, grid, blk, 0, st, L, U, t, p0, p1, r, 2.0 + am0, s, part, nshift,
                  g.bx, g.by, g.nbx);
}

void launch_ms_final(const Latt &L, MultiShiftState *s, int nshift, hipStream_t st)
{
    hipLaunchKernelGGL(k_ms_final, dim3(ew_blocks(2 * L.V)), dim3(kBlock), 0, st, s, nshift,
                       2 * L.V);
}

// ===========================================================================
// Forces and link updates
// ===========================================================================

// Fermion force z_mu at site x, returns Re(i z1), Re(i z2)  (src/QED2dDirac.jl:172-199)
__device__ __forceinline__ void ferm_force(const double2 *__restrict__ U,
                                           const double2 *__restrict__ X,
                                           const double2 *__restrict__ Y, const Latt &L, int i1,
                                           int i2, double &f1, double &f2)
{
    const long V = L.V;
    const int iu1 = (i1 + 1 == L.L1) ? 0 : i1 + 1;
    const int iu2 = (i2 + 1 == L.L2) ? 0 : i2 + 1;
    const long s = (long)i2 * L.L1 + i1;
    const long s1 = (long)i2 * L.L1 + iu1;
    const long s2 = (long)iu2 * L.L1 + i1;
    double2 U1 = U[s], U2 = U[V + s];
    if ((L.bc & 1) && i2 == L.L2 - 1) U2 = cneg<double>(U2);
    const double2 X1 = X[s], X2 = X[V + s], Y1 = Y[s], Y2 = Y[V + s];
    const double2 Xa1 = X[s1], Xa2 = X[V + s1], Ya1 = Y[s1], Ya2 = Y[V + s1];
    const double2 Xb1 = X[s2], Xb2 = X[V + s2], Yb1 = Y[s2], Yb2 = Y[V + s2];
    // z1 = U1* (X1(x+1)* (Y1+Y2) - X2(x+1)* (Y1+Y2)) - U1 (X1* + X2*)(Y1-Y2)(x+1)
    const double2 ys = cadd<double>(Y1, Y2);
    const double2 yd = csub<double>(Ya1, Ya2);
    double2 z1 = csub<double>(
        cmulc<double>(U1, csub<double>(cmulc<double>(Xa1, ys), cmulc<double>(Xa2, ys))),
        cmul<double>(U1, cadd<double>(cmulc<double>(X1, yd), cmulc<double>(X2, yd))));
    // z2 = U2* (X1(x+2)* (Y1 - i Y2) - X2(x+2)* (i Y1 + Y2))
    //      - U2 (X1* (Y1 + i Y2)(x+2) + X2* (i Y1 - Y2)(x+2))
    double2 z2 = csub<double>(
        cmulc<double>(U2, csub<double>(cmulc<double>(Xb1, csub<double>(Y1, cmuli<double>(Y2))),
                                       cmulc<double>(Xb2, cadd<double>(cmuli<double>(Y1), Y2)))),
        cmul<double>(U2, cadd<double>(cmulc<double>(X1, cadd<double>(Yb1, cmuli<double>(Yb2))),
                                      cmulc<double>(X2, csub<double>(cmuli<double>(Yb1), Yb2)))));
    // Re(i z) = -Im z
    f1 = -z1.y;
    f2 = -z2.y;
}

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
    hipLaunchKernelGGL(k_tr_dQwdU, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, U, X, Y, frc);
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
    hipLaunchKernelGGL(k_krnl_force, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, U, beta, frc1,
                       frc2);
}

// Gather form of the gauge force + fermion force + momentum update, one pass:
//   F(x,1) = -v(x) + v(x-2),  F(x,2) = v(x) - v(x-1),  v = beta Im P
// (the sums that src/QED2dAction.jl:47-50 scatters into frc1/frc2), then
//   mom += eps ((F_gauge) + cf F_fermion + facc)     (src/QED2dHMC.jl:478,554)
__global__ void k_force_mom(Latt L, const double2 *__restrict__ U, const double2 *__restrict__ X,
                            const double2 *__restrict__ Y, const double *__restrict__ facc,
                            double *__restrict__ mom, double beta, double eps, double cf)
{
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)gridDim.x * blockDim.x) {
        const int i2 = (int)(s / L.L1), i1 = (int)(s - (long)i2 * L.L1);
        const int id1 = (i1 == 0) ? L.L1 - 1 : i1 - 1;
        const int id2 = (i2 == 0) ? L.L2 - 1 : i2 - 1;
        const double v0 = beta * plaq_at(U, L, i1, i2).y;
        const double vd2 = beta * plaq_at(U, L, i1, id2).y;
        const double vd1 = beta * plaq_at(U, L, id1, i2).y;
        double g1 = -v0 + vd2;
        double g2 = v0 - vd1;
        if (X) {
            double f1, f2;
            ferm_force(U, X, Y, L, i1, i2, f1, f2);
            g1 = g1 + cf * f1;
            g2 = g2 + cf * f2;
        }
        if (facc) {
            g1 = g1 + facc[s];
            g2 = g2 + facc[L.V + s];
        }
        mom[s] = mom[s] + eps * g1;
        mom[L.V + s] = mom[L.V + s] + eps * g2;
    }
}
void launch_force_mom(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                      const double *facc, double *mom, double beta, double eps, double cf,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_force_mom, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, U, X, Y, facc,
                       mom, beta, eps, cf);
}

// facc (+)= rho * F_fermion  (src/QED2dHMC.jl:541-543)
__global__ void k_force_acc(Latt L, const double2 *__restrict__ U, const double2 *__restrict__ X,
                            const double2 *__restrict__ Y, double *__restrict__ facc, double rho,
                            int first)
{
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < L.V;
         s += (long)gridDim.x * blockDim.x) {
        const int i2 = (int)(s / L.L1), i1 = (int)(s - (long)i2 * L.L1);
        double f1, f2;
        ferm_force(U, X, Y, L, i1, i2, f1, f2);
        if (first) {
            facc[s] = rho * f1;
            facc[L.V + s] = rho * f2;
        } else {
            facc[s] = facc[s] + rho * f1;
            facc[L.V + s] = facc[L.V + s] + rho * f2;
        }
    }
}
void launch_force_acc(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                      double *facc, double rho, int first, hipStream_t st)
{
    hipLaunchKernelGGL(k_force_acc, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, U, X, Y, facc,
                       rho, first);
}

// U <- (cos(eps p), sin(eps p)) U  (src/QED2dAction.jl:55-64)
__global__ void k_updt_fld(double2 *__restrict__ U, const double *__restrict__ mom, double eps,
                           long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double sn, cs;
        sincos(eps * mom[i], &sn, &cs);
        U[i] = cmul<double>(mk<double>(cs, sn), U[i]);
    }
}
void launch_updt_fld(const Latt &L, double2 *U, const double *mom, double eps, hipStream_t st)
{
    hipLaunchKernelGGL(k_updt_fld, dim3(ew_blocks(2 * L.V)), dim3(kBlock), 0, st, U, mom, eps,
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
    hipLaunchKernelGGL(k_fill_momenta, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, mom, L.V, k0,
                       k1, traj);
}

__global__ void k_fill_pf(double2 *__restrict__ X, long V, uint32_t k0, uint32_t k1,
                          uint32_t traj, uint32_t flavour)
{
#pragma clang fp contract(off)
    const double sq2 = 1.4142135623730951;
    for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < V;
         s += (long)gridDim.x * blockDim.x) {
        for (uint32_t c = 0; c < 2; ++c) {
            double g1, g2;
            gauss_pair((uint32_t)s, c, kPurposePF + flavour, traj, k0, k1, &g1, &g2);
            X[c * V + s] = mk<double>(g1 / sq2, g2 / sq2);
        }
    }
}
void launch_fill_pseudofermion(const Latt &L, double2 *X, uint32_t k0, uint32_t k1,
                               uint32_t traj, uint32_t flavour, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_pf, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, X, L.V, k0, k1, traj,
                       flavour);
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
    hipLaunchKernelGGL(k_hot_start, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, U, L.V, k0, k1);
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
    hipLaunchKernelGGL(k_unfold, dim3(ew_blocks(L.V)), dim3(kBlock), 0, st, L, Uin, Uout,
                       reflect);
}

}  // namespace qed2d
