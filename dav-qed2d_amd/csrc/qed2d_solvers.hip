// qed2d_solvers.hip -- CG, mixed-precision CG and multi-shift CG on
// Q^2 + mu^2 (reference src/QED2dSolver.jl, src/QED2dRHMC.jl:261-290).
//
// One CG iteration k is two kernels, each a single pass of the marching
// Wilson stencil with the vector algebra fused into it:
//   K1(k): sum the |r_{k-1}|^2 partials of K2(k-1), convergence test,
//          p_k = r + beta p_{k-1} evaluated where it is read, t = Q p_k,
//          partials of |t|^2 and |p_k|^2
//   K2(k): alpha = |r_{k-1}|^2 / (|t|^2 + mu^2 |p_k|^2), Ap = Q t + mu^2 p_k,
//          x += alpha p_k, r -= alpha Ap, partials of |r_k|^2
// p^dag (Q^2 + mu^2) p = |Q p|^2 + mu^2 |p|^2 because Q is Hermitian, so the
// reference's dot(p, Ap) (src/QED2dSolver.jl:21) needs no pass of its own.
//
// Scalars never cross the grid inside a kernel: each kernel sums the
// previous kernel's per-block partials in block order (identical bits in
// every block), and block 0 of K1(k) commits what later kernels need into
// slot k & 1 of the state.  The search direction ping-pongs between two
// buffers (K1(k) reads P[(k-1) & 1], writes P[k & 1]) so the neighbours a
// block reads are never overwritten during the kernel.
#include <type_traits>

#include "qed2d_cg.h"

namespace qed2d {

static inline int ew_grid(long n)
{
    long b = (n + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

// ===========================================================================
// Start of a solve from x = 0 (src/QED2dSolver.jl:10-16)
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_init(Latt L, const double2 *__restrict__ b,
                                                    double2 *__restrict__ x,
                                                    c2<T> *__restrict__ r,
                                                    c2<T> *__restrict__ p0,
                                                    float2 *__restrict__ xs, double *Prr,
                                                    double *Pb, int bx, int by, int nbx,
                                                    int rows)
{
    MarchPos p = march_pos(L, bx, by, nbx, rows);
    double acc = 0.0;
    if (p.active) {
        for (int kk = 0; kk < rows; ++kk) {
            const int row = p.r0 + kk;
            if (row >= L.L2) break;
            for (int c = 0; c < 2; ++c) {
                const long i = c * L.V + (long)row * L.L1 + p.i1;
                const double2 v = b[i];
                x[i] = mk<double>(0.0, 0.0);
                r[i] = from_d2<T>(v);
                p0[i] = mk<T>(T(0), T(0));
                if (xs) xs[i] = mk<float>(0.0f, 0.0f);
                acc += v.x * v.x + v.y * v.y;
            }
        }
    }
    const double s = block_sum_t0(acc);
    if (threadIdx.x == 0) {
        Prr[p.blk] = s;
        Pb[p.blk] = s;
    }
}

// ---------------------------------------------------------------------------
// Folding.  A consumer kernel sums all partials of its producer in every one
// of its blocks, which costs O(blocks^2) reads.  Past kFoldMin producer blocks
// a one-block kernel folds them first, in place, to the single value the
// consumer would have computed: the same code (sum_prev_partials) over the
// same partials in the same order, so the bits do not depend on the folding.
// ---------------------------------------------------------------------------
constexpr int kFoldMin = kPartMax;  // consumers load at most this many partials
static_assert(kPartsUnfolded == kPartMax, "host and device partial limits differ");

__device__ __forceinline__ bool soft_ru_pending(const SolverState *S);

// pred: fold only while a soft reliable-update request is pending (the
// on-device update's residual partials; null: always)
__global__ void __launch_bounds__(kBlock) k_fold(double *part, int n, int stride,
                                                 const SolverState *pred)
{
    __shared__ double lds[kBlock / kWave + 1];
    if (pred && !soft_ru_pending(pred)) return;
    const double v0 = sum_prev_partials(part, n, stride, 0, lds);
    const double v1 = (stride > 1) ? sum_prev_partials(part, n, stride, 1, lds) : 0.0;
    if (threadIdx.x == 0) {
        part[0] = v0;
        if (stride > 1) part[1] = v1;
    }
}

// Returns the partial count a consumer must use.
static int fold(double *part, int n, int stride, hipStream_t st)
{
    if (n <= kFoldMin) return n;
    launch_k(k_fold, dim3(1), dim3(kBlock), 0, st, part, n, stride, (const SolverState *)nullptr);
    return 1;
}

void launch_cg_init_f64(const Latt &L, const double2 *b, double2 *x, double2 *r, double2 *p0,
                        CGPartials &P, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    launch_k(k_cg_init<double>, dim3(g.nblocks()), dim3(kBlock), 0, st, L, b, x, r, p0,
                       (float2 *)nullptr, P.rr[0], P.b, g.bx, g.by, g.nbx, rows);
    P.nrr[0] = fold(P.rr[0], g.nblocks(), 1, st);
    P.nb = fold(P.b, g.nblocks(), 1, st);
}

void launch_cg_init_mixed(const Latt &L, const double2 *b, double2 *x, float2 *rs, float2 *xs,
                          float2 *p0, CGPartials &P, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    launch_k(k_cg_init<float>, dim3(g.nblocks()), dim3(kBlock), 0, st, L, b, x, rs, p0,
                       xs, P.rr[0], P.b, g.bx, g.by, g.nbx, rows);
    P.nrr[0] = fold(P.rr[0], g.nblocks(), 1, st);
    P.nb = fold(P.b, g.nblocks(), 1, st);
}

// ===========================================================================
// Residual of an FP64 x: r = b - (Q t + mu^2 x), t = Q x  (the guess branch
// of src/QED2dSolver.jl:65-72, and the reliable update of the mixed solver)
// ===========================================================================
// Mode 3 is the on-device reliable update queued at a chunk boundary: it runs
// only while a soft request is pending on a running solve (ru_req && !stop_k,
// which none of the update's kernels modify), and pads its |r|^2 partials with
// zeros up to npad, the block count of the K2 whose slot it overwrites (adding
// +0.0 leaves the consumer's sum bit for bit as over the real partials).
// Mode 4 is the update queued behind the chunk predicted to end a solve: it
// runs only if K1(k) paused there (need_ru == k, which the K1 that runs next
// clears), pads like mode 3 and resumes the solve at K1(k) like mode 2.
__device__ __forceinline__ bool soft_ru_pending(const SolverState *S)
{
    return S->ru_req != 0 && S->stop_k == 0;
}

template <int R>
__global__ void __launch_bounds__(kBlock) k_cg_resid(Latt L, const double2 *__restrict__ U,
                                                     const double2 *__restrict__ t,
                                                     EpiResid e0, double mass2, SolverState *S,
                                                     double *Prr, double *Pb, int k, int mode,
                                                     int npad, int bx, int by, int nbx)
{
    if (mode == 3 && !soft_ru_pending(S)) return;
    if (mode == 4 && S->need_ru != k) return;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    EpiResid epi = e0;
    epi.mu2 = S->mu2;
    epi.rr = 0.0;
    epi.bb = 0.0;
    if (p.active) {
        SrcPlain<double> src{t, L.V};
        q_march<double, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a, bb;
    block_sum2_t0(epi.rr, epi.bb, a, bb);
    if (threadIdx.x == 0) {
        Prr[p.blk] = a;
        if (mode < 2) Pb[p.blk] = bb;
    }
    if (p.blk == 0 && threadIdx.x == 0) {
        for (int i = (int)gridDim.x; i < npad; ++i) Prr[i] = 0.0;
        S->true_at = k;  // K1(k) will see an FP64 true residual
        if (mode >= 2) S->n_ru += 1;
        if (mode == 2) {
            S->need_ru = 0;
            S->stop_k = 0;  // resume: K1(k) runs again
        }
        if (mode == 4) S->stop_k = 0;  // resume (need_ru is left to K1(k))
    }
}

void launch_cg_resid(const Latt &L, const double2 *U, const double2 *b, const double2 *t,
                     const double2 *x, double2 *r, float2 *rs, double2 *p0, float2 *p0s,
                     double am0, SolverState *s, CGPartials &P, int k, int mode, int rows,
                     hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    EpiResid e{b, x, r, rs, mode == 0 ? p0 : nullptr, mode == 1 ? p0s : nullptr, 0.0, L.V, 0.0,
               0.0};
    dim3 grid(g.nblocks()), blk(kBlock);
    // mode 3 keeps the slot's count, that of the K2 before it: a folded count
    // (1) gets the residual's partials folded the same way, an unfolded one
    // (which the caller checked is at least this grid's) zero padding up to it
    const int nslot = P.nrr[(k - 1) & 1];
    const bool fold3 = (mode == 3 && nslot == 1 && g.nblocks() > 1);
    const int npad = ((mode == 3 && !fold3) || mode == 4) ? nslot : 0;
    DISPATCH_ROWS(rows, k_cg_resid, grid, blk, 0, st, L, U, t, e, 2.0 + am0, s, P.rr[(k - 1) & 1],
                  P.b, k, mode, npad, g.bx, g.by, g.nbx);
    if (mode == 4) return;  // (the caller made sure nslot >= this grid, unfolded)
    if (mode == 3) {
        if (fold3)
            launch_k(k_fold, dim3(1), dim3(kBlock), 0, st, P.rr[(k - 1) & 1], g.nblocks(), 1,
                     (const SolverState *)s);
        return;
    }
    P.nrr[(k - 1) & 1] = fold(P.rr[(k - 1) & 1], g.nblocks(), 1, st);
    if (mode != 2) P.nb = fold(P.b, g.nblocks(), 1, st);
}

// Predicated FP64 apply of the on-device reliable update: so = Q si while a
// soft request is pending (see soft_ru_pending).
// (pk > 0: instead while K1(pk) is paused for an update, need_ru == pk;
// pk < 0: once the solve has converged, status == 1)
template <int R>
__global__ void __launch_bounds__(kBlock) k_q_apply_ru(Latt L, const double2 *__restrict__ U,
                                                       const double2 *__restrict__ si,
                                                       double2 *__restrict__ so, double mass2,
                                                       const SolverState *S, int bx, int by,
                                                       int nbx, int pk)
{
    if (pk > 0 ? S->need_ru != pk : pk < 0 ? S->status != 1 : !soft_ru_pending(S)) return;
    MarchPos p = march_pos(L, bx, by, nbx, R);
    if (!p.active) return;
    SrcPlain<double> src{si, L.V};
    EpiStore<double> epi{so, L.V};
    q_march<double, R>(L, U, mass2, src, epi, p.i1, p.r0);
}
void launch_q_apply_ru(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                       double am0, const SolverState *S, int rows, hipStream_t st, int pk)
{
    MarchGrid g = march_grid(L, rows);
    DISPATCH_ROWS(rows, k_q_apply_ru, dim3(g.nblocks()), dim3(kBlock), 0, st, L, U, si, so,
                  2.0 + am0, S, g.bx, g.by, g.nbx, pk);
}
int resid_blocks(const Latt &L, int rows) { return march_grid(L, rows).nblocks(); }
int cg2_f32_blocks(const Latt &L, int rows)
{
    return pair_ok(L) ? march_grid2(L, rows).nblocks() : march_grid(L, rows).nblocks();
}

// x += double(xs); xs = 0 (with the pending xs updates first when P0 is
// given, before K1(k), k = stop_k: that of p_{k-2} if k is even and K2(k-2)
// ran after the last true residual, then that of p_{k-1}, see
// SolverState::alpha_xk; kb > 0: the on-device update before K1(kb), k = kb,
// run only while a soft request is pending)
__global__ void k_ru_accum(double2 *__restrict__ x, float2 *__restrict__ xs,
                           const float2 *__restrict__ P0, const float2 *__restrict__ P1,
                           const SolverState *S, long n, int kb, int pk)
{
    if (pk > 0) {
        if (S->need_ru != pk) return;
    } else if (kb > 0 && !soft_ru_pending(S)) {
        return;
    }
    const float2 *pp = nullptr, *pq = nullptr;
    float a = 0.0f, a2 = 0.0f;
    if (P0) {
        const int k = pk > 0 ? pk : kb > 0 ? kb : S->stop_k;
        pp = ((k - 1) & 1) ? P1 : P0;
        a = (float)S->alpha_x;
        const int t0 = S->true_at > 1 ? S->true_at : 1;
        if ((k & 1) == 0 && k - 2 >= t0) {
            pq = (k & 1) ? P1 : P0;
            a2 = (float)S->alpha_xk[k & 1];
        }
    }
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        float2 d = xs[i];
        if (pq) d = xpa(d, a2, pq[i]);
        if (pp) d = xpa(d, a, pp[i]);
        const double2 v = x[i];
        x[i] = mk<double>(v.x + (double)d.x, v.y + (double)d.y);
        xs[i] = mk<float>(0.0f, 0.0f);
    }
}
void launch_ru_accum(const Latt &L, double2 *x, float2 *xs, const float2 *p0, const float2 *p1,
                     const SolverState *S, hipStream_t st, int kb, int pk)
{
    launch_k(k_ru_accum, dim3(ew_grid(2 * L.V)), dim3(kBlock), 0, st, x, xs, p0, p1, S,
                       2 * L.V, kb, pk);
}

__global__ void k_to_f32(const double2 *__restrict__ in, float2 *__restrict__ out, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        out[i] = from_d2<float>(in[i]);
}
void launch_to_f32(const double2 *in, float2 *out, long n, hipStream_t st)
{
    launch_k(k_to_f32, dim3(ew_grid(n)), dim3(kBlock), 0, st, in, out, n);
}

// out = in widened to FP64 (exact)
__global__ void k_to_f64(const float2 *__restrict__ in, double2 *__restrict__ out, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        out[i] = mk<double>((double)in[i].x, (double)in[i].y);
}
void launch_to_f64(const float2 *in, double2 *out, long n, hipStream_t st)
{
    launch_k(k_to_f64, dim3(ew_grid(n)), dim3(kBlock), 0, st, in, out, n);
}

// ===========================================================================
// CG iteration kernels
// ===========================================================================
// K1(k).  FP64 (T = double): the reference's test |r|^2 < eps |b|^2.  FP32
// (T = float, the iterations of the mixed solver): the iterated residual asks
// for a reliable update once it has dropped by delta since the last one, looks
// converged, or the iteration limit is reached; only an FP64 true residual
// (true_at == k) ends the solve.
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_cg1(Latt L, const c2<T> *__restrict__ U,
                                                const c2<T> *__restrict__ r, c2<T> *P0,
                                                c2<T> *P1, c2<T> *__restrict__ t, T mass2,
                                                SolverState *S, double *Ptt, const double *Prr,
                                                const double *Pb, int nrr, int nbb, int k, int bx,
                                                int by, int nbx, int soft)
{
    __shared__ double lds[kBlock / kWave + 1];
    // The solver state's loads go out first, the partials' next, the first
    // rows' last; nothing waits for them before the rows' loads are in flight.
    // A solve that stopped before K1(k) returns before any store.
    const SolverState *Sv = state_v(S);
    const int sk = Sv->stop_k;
    const K1In in = k1_in(Sv, k);
    const int w = k & 1, rd = (k - 1) & 1;
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    PartialsAhead prr;
    prr.load(Prr, nrr, 1, 0);
    const c2<T> *pold = rd ? P1 : P0;
    c2<T> *pnew = w ? P1 : P0;
    EpiK1<T> epi{pnew, t, L.V, 0.0, 0.0};
    SrcXpay<T> src{r, pold, T(0), L.V};
    auto ready = [&]() -> bool {
        if (sk != 0 && sk < k) return false;
        double beta;
        if (!cg_k1_decide<std::is_same<T, float>::value>(S, in, prr.sum(), Pb, nbb, k, p.blk,
                                                          lds, beta, nullptr, soft != 0))
            return false;
        src.beta = (T)beta;
        return true;
    };
    if (!q_march_r<T, R>(L, U, mass2, src, epi, p.i1, p.r0, p.active, ready)) return;
    double a, b;
    block_sum2_t0(epi.tt, epi.pp, a, b);
    if (threadIdx.x == 0) {
        Ptt[2 * p.blk] = a;
        Ptt[2 * p.blk + 1] = b;
    }
}

template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_cg2(Latt L, const c2<T> *__restrict__ U,
                                                const c2<T> *__restrict__ t, const c2<T> *P0,
                                                const c2<T> *P1, c2<T> *__restrict__ x,
                                                c2<T> *__restrict__ r, T mass2, SolverState *S,
                                                const double *Ptt, double *Prr, int ntt, int k,
                                                int bx, int by, int nbx, StateSnap snap)
{
    // The solver state's loads go out first, the partials' next, the first
    // rows' last; nothing waits for them before the rows' loads are in flight.
    // A solve that stopped before K2(k) returns before any store, and alpha
    // (src/QED2dSolver.jl:21-22) is formed once all are in.
    const int w = k & 1;
    const SolverState *Sv = state_v(S);
    const int sk = Sv->stop_k;
    const double nw = Sv->norm[w], mu2 = Sv->mu2;
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    PartialsAhead2 ptt;
    ptt.load(Ptt, ntt);
    const c2<T> *pv = w ? P1 : P0;
    EpiK2<T> epi{pv, x, r, T(0), T(0), L.V, 0.0};
    auto ready = [&]() -> bool {
        if (sk != 0 && sk <= k) return false;
        double tt, pp;
        ptt.sum(tt, pp);
        epi.alpha = (T)(nw / (tt + mu2 * pp));
        epi.mu2 = (T)mu2;
        return true;
    };
    SrcPlain<T> src{t, L.V};
    if (q_march_r<T, R>(L, U, mass2, src, epi, p.i1, p.r0, p.active, ready)) {
        const double a = block_sum_t0(epi.rr);
        if (threadIdx.x == 0) Prr[p.blk] = a;
    }
    if (p.blk == 0 && threadIdx.x == 0) put_snap(snap, S);
}

// ---- FP32 pair kernels (two sites per thread along i1) -------------------
// G: K1(k) queued again behind a mode-4 reliable update, run only if that
// update ran (K1(k) had paused and was resumed); a separate instance, so the
// check costs the ordinary K1 nothing.  F: k is even, the K1 that applies the
// pending x updates (SolverState::alpha_xk); the K1 of an odd k neither reads
// nor writes x.
template <int R, bool PK, bool G, bool F>
__global__ void __launch_bounds__(kBlock) k_cg1v(Latt L, LinksF32 U,
                                                 const float2 *__restrict__ r, float2 *P0,
                                                 float2 *P1, float2 *__restrict__ t,
                                                 float2 *__restrict__ x, float mass2,
                                                 SolverState *S, double *Ptt, const double *Prr,
                                                 const double *Pb, int nrr, int nbb, int k, int bx,
                                                 int by, int nbx, int soft)
{
    __shared__ double lds[kBlock / kWave + 1];
    // The solver state's loads go out first, the partials' next, the first
    // rows' last (x with them whenever an update may be pending); nothing waits
    // for them before the rows' loads are in flight.  A solve that stopped
    // before K1(k) returns before any store.
    const SolverState *Sv = state_v(S);
    const int sk = Sv->stop_k;
    const K1In in = k1_in(Sv, k);
    const float ax = (float)Sv->alpha_x;
    const int w = k & 1, rd = (k - 1) & 1;
    // (only an even k applies the pending x updates, see SolverState::alpha_xk)
    constexpr bool flush = F;
    const float ax2 = flush ? (float)Sv->alpha_xk[w] : 0.0f;
    const MarchPos p = march_pos2(L, bx, by, nbx, R);
    PartialsAhead prr;
    prr.load(Prr, nrr, 1, 0);
    float2 *pold = rd ? P1 : P0;
    EpiK1v epi{w ? P1 : P0, t, x, 0.0f, false, flush, L.V, 0.0, 0.0};
    epi.ldp2 = flush;
    SrcXpay2 src{r, pold, 0.0f, L.V};
    int gate = 1;
    if constexpr (G) gate = (Sv->need_ru == k && in.true_at == k) ? 1 : 0;
    auto ready = [&]() -> bool {
        if (sk != 0 && sk < k) return false;
        if (G && !gate) return false;
        double beta;
        bool fresh;
        if (!cg_k1_decide<true>(S, in, prr.sum(), Pb, nbb, k, p.blk, lds, beta, &fresh,
                                soft != 0))
            return false;
        src.beta = (float)beta;
        epi.alpha = ax;
        epi.alpha2 = ax2;
        epi.upd = flush && !fresh;
        // p_{k-2}'s update is pending if K2(k-2) ran after the last true
        // residual (K1(true_at), or K1(1))
        epi.two = k - 2 >= (in.true_at > 1 ? in.true_at : 1);
        return true;
    };
    if (!q_march2t<R, PK>(L, U, mass2, src, epi, p.i1, p.r0, p.active, ready)) return;
    double a, b;
    block_sum2_t0(epi.tt, epi.pp, a, b);
    if (threadIdx.x == 0) {
        Ptt[2 * p.blk] = a;
        Ptt[2 * p.blk + 1] = b;
    }
}

template <int R, bool MU, bool PK>
__device__ __forceinline__ void cg2v_body(const Latt &L, const LinksF32 &U,
                                          const float2 *__restrict__ t, const float2 *P0,
                                          const float2 *P1, float2 *__restrict__ r, float mass2,
                                          SolverState *S, const double *Ptt, double *Prr, int ntt,
                                          int k, int bx, int by, int nbx, int guard,
                                          const StateSnap &snap)
{
    // The solver state's loads go out first, the partials' next, the first
    // rows' last; nothing waits for them before the rows' loads are in flight.
    // A solve that stopped before K2(k) returns before any store.
    const int w = k & 1;
    const SolverState *Sv = state_v(S);
    const int sk = Sv->stop_k, ta = Sv->true_at;
    const double nw = Sv->norm[w], mu2 = Sv->mu2;
    const MarchPos p = march_pos2(L, bx, by, nbx, R);
    PartialsAhead2 ptt;
    ptt.load(Ptt, ntt);
    EpiK2v<MU> epi{w ? P1 : P0, r, 0.0f, 0.0f, L.V, 0.0};
    double alpha = 0.0;
    auto ready = [&]() -> bool {
        if (sk != 0 && sk <= k) return false;
        // (guard: K2(k) queued again behind a mode-4 reliable update runs only
        // if that update ran, which made K1(k) see a true residual)
        if (guard && ta != k) return false;
        double tt, pp;
        ptt.sum(tt, pp);
        alpha = nw / (tt + mu2 * pp);
        epi.alpha = (float)alpha;
        epi.mu2 = (float)mu2;
        return true;
    };
    SrcPlain2 src{t, L.V};  // alpha: src/QED2dSolver.jl:21-22
    if (q_march2t<R, PK>(L, U, mass2, src, epi, p.i1, p.r0, p.active, ready)) {
        const double a = block_sum_t0(epi.rr);
        if (threadIdx.x == 0) {
            Prr[p.blk] = a;
            if (p.blk == 0) {
                S->alpha_x = alpha;  // for K1(k + 1)
                S->alpha_xk[k & 1] = alpha;  // for K1(k + 1), K1(k + 2)
            }
        }
    }
    if (p.blk == 0 && threadIdx.x == 0) put_snap(snap, S);
}
#define CG2V_KERNEL(NAME, MU)                                                              \
    template <int R, bool PK>                                                              \
    __global__ void __launch_bounds__(kBlock)                                              \
        NAME(Latt L, LinksF32 U, const float2 *__restrict__ t,                            \
             const float2 *P0, const float2 *P1, float2 *__restrict__ r, float mass2,      \
             SolverState *S, const double *Ptt, double *Prr, int ntt, int k, int bx, int by, \
             int nbx, int guard, StateSnap snap)                                           \
    {                                                                                      \
        cg2v_body<R, MU, PK>(L, U, t, P0, P1, r, mass2, S, Ptt, Prr, ntt, k, bx, by, nbx,  \
                             guard, snap);                                                 \
    }
CG2V_KERNEL(k_cg2v, false)
CG2V_KERNEL(k_cg2vm, true)  // mu2 != 0

// k_cg1v<R, PK, G, F> by rows
template <bool PK, bool G, bool F, typename... A>
static void k1v_rows(int rows, dim3 grid, dim3 blk, hipStream_t st, A... args)
{
    switch (rows) {
    case 1: launch_k((k_cg1v<1, PK, G, F>), grid, blk, 0, st, args...); break;
    case 2: launch_k((k_cg1v<2, PK, G, F>), grid, blk, 0, st, args...); break;
    case 4: launch_k((k_cg1v<4, PK, G, F>), grid, blk, 0, st, args...); break;
    case 8: launch_k((k_cg1v<8, PK, G, F>), grid, blk, 0, st, args...); break;
    default: launch_k((k_cg1v<16, PK, G, F>), grid, blk, 0, st, args...); break;
    }
}
template <bool PK, bool G, typename... A>
static void k1v_launch(bool even, int rows, dim3 grid, dim3 blk, hipStream_t st, A... args)
{
    if (even)
        k1v_rows<PK, G, true>(rows, grid, blk, st, args...);
    else
        k1v_rows<PK, G, false>(rows, grid, blk, st, args...);
}

template <typename T>
static void cg1_impl(const Latt &L, const c2<T> *U, const c2<T> *r, c2<T> *p0, c2<T> *p1,
                     c2<T> *t, c2<T> *x, T am0, SolverState *S, CGPartials &P, int k, int rows,
                     hipStream_t st, int packed = 0, int soft = 0, bool guard = false)
{
    const T m2 = T(2) + am0;
    double *ptt = P.tt[k & 1];
    const double *prr = P.rr[(k - 1) & 1];
    const int nrr = P.nrr[(k - 1) & 1], nbb = P.nb;
    if constexpr (std::is_same<T, float>::value) {
        if (pair_ok(L)) {
            MarchGrid g = march_grid2(L, rows);
            dim3 grid(g.nblocks()), blk(kBlock);
            const LinksF32 lk{U, packed};
            const bool even = (k & 1) == 0;
#define K1V_ARGS L, lk, r, p0, p1, t, x, m2, S, ptt, prr, (const double *)P.b, nrr, nbb, k, g.bx, \
                 g.by, g.nbx, soft
            if (packed && guard)
                k1v_launch<true, true>(even, rows, grid, blk, st, K1V_ARGS);
            else if (packed)
                k1v_launch<true, false>(even, rows, grid, blk, st, K1V_ARGS);
            else if (guard)
                k1v_launch<false, true>(even, rows, grid, blk, st, K1V_ARGS);
            else
                k1v_launch<false, false>(even, rows, grid, blk, st, K1V_ARGS);
#undef K1V_ARGS
            P.ntt[k & 1] = fold(ptt, g.nblocks(), 2, st);
            return;
        }
    }
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    switch (rows) {
    case 1: launch_k((k_cg1<T, 1>), grid, blk, 0, st, L, U, r, p0, p1, t, m2, S, ptt, prr, P.b, nrr, nbb, k, g.bx, g.by, g.nbx, soft); break;
    case 2: launch_k((k_cg1<T, 2>), grid, blk, 0, st, L, U, r, p0, p1, t, m2, S, ptt, prr, P.b, nrr, nbb, k, g.bx, g.by, g.nbx, soft); break;
    case 4: launch_k((k_cg1<T, 4>), grid, blk, 0, st, L, U, r, p0, p1, t, m2, S, ptt, prr, P.b, nrr, nbb, k, g.bx, g.by, g.nbx, soft); break;
    case 8: launch_k((k_cg1<T, 8>), grid, blk, 0, st, L, U, r, p0, p1, t, m2, S, ptt, prr, P.b, nrr, nbb, k, g.bx, g.by, g.nbx, soft); break;
    default: launch_k((k_cg1<T, 16>), grid, blk, 0, st, L, U, r, p0, p1, t, m2, S, ptt, prr, P.b, nrr, nbb, k, g.bx, g.by, g.nbx, soft); break;
    }
    P.ntt[k & 1] = fold(ptt, g.nblocks(), 2, st);
}

template <typename T>
static void cg2_impl(const Latt &L, const c2<T> *U, const c2<T> *t, const c2<T> *p0,
                     const c2<T> *p1, c2<T> *x, c2<T> *r, T am0, SolverState *S,
                     CGPartials &P, int k, int rows, bool mu, hipStream_t st, int packed = 0,
                     int guard = 0, StateSnap snap = StateSnap{nullptr, nullptr, 0})
{
    const T m2 = T(2) + am0;
    const double *ptt = P.tt[k & 1];
    double *prr = P.rr[k & 1];
    const int ntt = P.ntt[k & 1];
    if constexpr (std::is_same<T, float>::value) {
        if (pair_ok(L)) {
            MarchGrid g = march_grid2(L, rows);
            dim3 grid(g.nblocks()), blk(kBlock);
            // x is updated by the next K1 (or the reliable update's flush)
            const LinksF32 lk{U, packed};
            if (mu) {
                DISPATCH_ROWS_PK(rows, packed, k_cg2vm, grid, blk, 0, st, L, lk, t, p0, p1, r, m2,
                                 S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, guard, snap);
            } else {
                DISPATCH_ROWS_PK(rows, packed, k_cg2v, grid, blk, 0, st, L, lk, t, p0, p1, r, m2,
                                 S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, guard, snap);
            }
            P.nrr[k & 1] = fold(prr, g.nblocks(), 1, st);
            return;
        }
    }
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    switch (rows) {
    case 1: launch_k((k_cg2<T, 1>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, snap); break;
    case 2: launch_k((k_cg2<T, 2>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, snap); break;
    case 4: launch_k((k_cg2<T, 4>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, snap); break;
    case 8: launch_k((k_cg2<T, 8>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, snap); break;
    default: launch_k((k_cg2<T, 16>), grid, blk, 0, st, L, U, t, p0, p1, x, r, m2, S, ptt, prr, ntt, k, g.bx, g.by, g.nbx, snap); break;
    }
    P.nrr[k & 1] = fold(prr, g.nblocks(), 1, st);
}

void launch_cg_k1_f64(const Latt &L, const double2 *U, const double2 *r, double2 *p0,
                      double2 *p1, double2 *t, double am0, SolverState *s, CGPartials &P,
                      int k, int rows, hipStream_t st)
{
    cg1_impl<double>(L, U, r, p0, p1, t, nullptr, am0, s, P, k, rows, st);
}
void launch_cg_k2_f64(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                      const double2 *p1, double2 *x, double2 *r, double am0, SolverState *s,
                      CGPartials &P, int k, int rows, hipStream_t st, StateSnap snap)
{
    cg2_impl<double>(L, U, t, p0, p1, x, r, am0, s, P, k, rows, false, st, 0, 0, snap);
}
void launch_cg_k1_f32(const Latt &L, const void *U, bool packed, const float2 *r, float2 *p0,
                      float2 *p1, float2 *t, float2 *x, float am0, SolverState *s, CGPartials &P,
                      int k, int rows, hipStream_t st, bool soft, bool guard)
{
    cg1_impl<float>(L, (const float2 *)U, r, p0, p1, t, x, am0, s, P, k, rows, st, packed,
                    soft ? 1 : 0, guard);
}
void launch_cg_k2_f32(const Latt &L, const void *U, bool packed, const float2 *t,
                      const float2 *p0, const float2 *p1, float2 *x, float2 *r, float am0,
                      bool mu, SolverState *s, CGPartials &P, int k, int rows, hipStream_t st,
                      bool guard, StateSnap snap)
{
    cg2_impl<float>(L, (const float2 *)U, t, p0, p1, x, r, am0, s, P, k, rows, mu, st, packed,
                    guard ? 1 : 0, snap);
}

// U(1) links packed into one float each (the format link_dec reads): the
// smaller-magnitude component, rounded to a float whose last mantissa bit is
// 0 (towards the exact value), bit 30 set when it is Re, the last bit set when
// the other component is negative.  The format assumes unit modulus: the
// larger component comes back as sqrt(1 - v^2), off by about 0.7 ||U|^2 - 1|.
// Links with ||U|^2 - 1| > 2^-23 (FP32 epsilon) set S->bad_links, so a packed
// link is never further from its FP64 value than FP32 rounding of its two
// components (~1e-7); other links make the solver use FP32 complex links.
__global__ void k_link_pack(const double2 *__restrict__ in, float *__restrict__ out, long n,
                            SolverState *S)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double2 u = in[i];
        const bool sw = fabs(u.y) > fabs(u.x);  // store Re
        const double v = sw ? u.x : u.y, o = sw ? u.y : u.x;
        const float w = (float)v;
        uint32_t b = __float_as_uint(w);
        if (b & 1u) b = (fabs(v) > fabs((double)w)) ? b + 1u : b - 1u;
        if (sw) b |= 0x40000000u;
        if (o < 0.0) b |= 1u;
        out[i] = __uint_as_float(b);
        if (fabs(u.x * u.x + u.y * u.y - 1.0) > 0x1p-23) S->bad_links = 1;
    }
}
void launch_link_pack(const double2 *U, float *out, long n, SolverState *S, hipStream_t st)
{
    launch_k(k_link_pack, dim3(ew_grid(n)), dim3(kBlock), 0, st, U, out, n, S);
}

// A copy of the solver state into pinned host memory for the host's chunk
// planning, then the sequence number v into *seq with a release at system
// scope: the host that reads v there finds the copy complete.  (No blit and no
// event between the chunks: the next chunk's kernels follow at once.)
__global__ void k_state_snap(const SolverState *__restrict__ S, SolverState *out, int *seq, int v)
{
    if (threadIdx.x == 0) put_snap(StateSnap{out, seq, v}, S);
}
void launch_state_snap(const SolverState *S, SolverState *out, int *seq, int v, hipStream_t st)
{
    launch_k(k_state_snap, dim3(1), dim3(kWave), 0, st, S, out, seq, v);
}

// ===========================================================================
// Multi-shift CG (Jegerlehner, hep-lat/9612014).  The systems
// (Q^2 + sigma_i) x_i = b share the Krylov space of the base system
// (Q^2 + sigma_b), sigma_b the smallest shift, whose residual r_j fixes every
// shifted residual up to a scalar, r^i_j = zeta^i_j r_j.  With
// s = sigma_i - sigma_b and p_j = r_j + beta_j p_{j-1}:
//   zeta_j    = zeta_{j-1} zeta_{j-2} alpha_{j-2} /
//               (alpha_{j-1} beta_{j-1} (zeta_{j-2} - zeta_{j-1})
//                + zeta_{j-2} alpha_{j-2} (1 + s alpha_{j-1}))
//   alpha^i_{j-1} = alpha_{j-1} zeta_j / zeta_{j-1},  beta^i_j = beta_j (zeta_j / zeta_{j-1})^2
//   x_i += alpha^i_{j-1} p^i_{j-1},  p^i_j = zeta_j r_j + beta^i_j p^i_{j-1}
// (zeta_0 = zeta_{-1} = alpha_{-1} = 1, beta_0 = 0).  K1(k) handles j = k - 1
// and fuses every shift's x and p update into the stencil pass; a shift stops
// once zeta_j^2 |r_j|^2 < eps |b|^2.  Bytes per site and iteration (FP64):
// K1 160 + 128 per active shift, K2 160.
// ===========================================================================
template <int R>
__global__ void __launch_bounds__(kBlock) k_ms1(Latt L, const double2 *__restrict__ U,
                                                const double2 *__restrict__ r, double2 *P0,
                                                double2 *P1, double2 *__restrict__ t,
                                                double mass2, MultiShiftState *S, double *Ptt_w,
                                                const double *Ptt_rd, const double *Prr_rd,
                                                int nrr, int ntt, int k, int nshift, int bx,
                                                int by, int nbx)
{
    const int sk = S->stop_k;
    if (sk != 0 && sk < k) return;
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    ms1_body<double, R>(L, U, r, P0, P1, t, mass2, S, Ptt_w, Ptt_rd, Prr_rd, nrr, ntt, k, nshift,
                        p);
}

template <int R>
__global__ void __launch_bounds__(kBlock) k_ms2(Latt L, const double2 *__restrict__ U,
                                                const double2 *__restrict__ t, const double2 *P0,
                                                const double2 *P1, double2 *__restrict__ r,
                                                double mass2, MultiShiftState *S,
                                                const double *Ptt_w, double *Prr_w, int ntt, int k,
                                                int bx, int by, int nbx)
{
    const int sk = S->stop_k;
    if (sk != 0 && sk <= k) return;
    const MarchPos p = march_pos(L, bx, by, nbx, R);
    ms2_body<double, R>(L, U, t, P0, P1, r, mass2, S, Ptt_w, Prr_w, ntt, k, p);
}

// The shift updates of step k, after K1(k) committed their coefficients: grid
// x over the n spinor elements, y over the shifts.  A step that K1 skipped (the
// solve stopped earlier) updates nothing.
__global__ void __launch_bounds__(kBlock) k_ms_shifts(const MultiShiftState *S,
                                                      const double2 *__restrict__ r, long n, int k)
{
    const int sk = S->stop_k;
    if (sk != 0 && sk < k) return;
    const long i = blockIdx.x * (long)kBlock + threadIdx.x;
    if (i >= n) return;
    ms_shift_update(S, r, i, blockIdx.y);
}
void launch_ms_shifts(const MultiShiftState *s, const double2 *r, long n, int k, int nshift,
                      hipStream_t st)
{
    launch_k(k_ms_shifts, dim3((unsigned)((n + kBlock - 1) / kBlock), nshift), dim3(kBlock), 0,
             st, s, r, n, k);
}

// r = b, p0 = 0, x_i = 0, p_i = 0, |b|^2 partials.
__global__ void __launch_bounds__(kBlock) k_ms_init(Latt L, const double2 *__restrict__ b,
                                                    double2 *__restrict__ r,
                                                    double2 *__restrict__ p0,
                                                    const MultiShiftState *S, double *Prr,
                                                    int nshift, int bx, int by, int nbx, int rows)
{
    MarchPos p = march_pos(L, bx, by, nbx, rows);
    double acc = 0.0;
    const double2 z = mk<double>(0.0, 0.0);
    if (p.active) {
        for (int kk = 0; kk < rows; ++kk) {
            const int row = p.r0 + kk;
            if (row >= L.L2) break;
            for (int c = 0; c < 2; ++c) {
                const long i = c * L.V + (long)row * L.L1 + p.i1;
                const double2 v = b[i];
                r[i] = v;
                p0[i] = z;
                for (int k = 0; k < nshift; ++k) {
                    ms_vec<double>(S->x[k])[i] = z;
                    ms_vec<double>(S->ps[k])[i] = z;
                }
                acc += v.x * v.x + v.y * v.y;
            }
        }
    }
    const double s = block_sum_t0(acc);
    if (threadIdx.x == 0) Prr[p.blk] = s;
}

void launch_ms_init(const Latt &L, const double2 *b, double2 *r, double2 *p0, MultiShiftState *s,
                    CGPartials &P, int nshift, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    launch_k(k_ms_init, dim3(g.nblocks()), dim3(kBlock), 0, st, L, b, r, p0, s, P.rr[0],
                       nshift, g.bx, g.by, g.nbx, rows);
    P.nrr[0] = fold(P.rr[0], g.nblocks(), 1, st);
}

void launch_ms_k1(const Latt &L, const double2 *U, const double2 *r, double2 *p0, double2 *p1,
                  double2 *t, double am0, MultiShiftState *s, CGPartials &P, int k,
                  int nshift, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    DISPATCH_ROWS(rows, k_ms1, grid, blk, 0, st, L, U, r, p0, p1, t, 2.0 + am0, s, P.tt[k & 1],
                  P.tt[(k - 1) & 1], P.rr[(k - 1) & 1], P.nrr[(k - 1) & 1], P.ntt[(k - 1) & 1], k,
                  nshift, g.bx, g.by, g.nbx);
    P.ntt[k & 1] = fold(P.tt[k & 1], g.nblocks(), 2, st);
}

void launch_ms_k2(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                  const double2 *p1, double2 *r, double am0, MultiShiftState *s,
                  CGPartials &P, int k, int nshift, int rows, hipStream_t st)
{
    MarchGrid g = march_grid(L, rows);
    dim3 grid(g.nblocks()), blk(kBlock);
    DISPATCH_ROWS(rows, k_ms2, grid, blk, 0, st, L, U, t, p0, p1, r, 2.0 + am0, s, P.tt[k & 1],
                  P.rr[k & 1], P.ntt[k & 1], k, g.bx, g.by, g.nbx);
    P.nrr[k & 1] = fold(P.rr[k & 1], g.nblocks(), 1, st);
}

}  // namespace qed2d
