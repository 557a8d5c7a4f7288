// qed2d_batch.hip -- batched ensembles: many independent lattices ("chains")
// of different sizes, couplings and masses advanced together.  Each kernel
// here is one launch over every chain: a block finds its chain by binary
// search over the launch's block offsets (tile kernels) or site offsets
// (site kernels), then runs the same device code as the single-lattice
// kernels on that chain's slice of the arena.  A chain tiled like a
// single-lattice launch of the same tile height therefore produces the same
// bits as that launch, partial sums included.
//
// Per-chain CG: every chain keeps its own iteration index in its solver state
// (K1 reads kk, commits k2 for K2; K2 advances kk), so chains that finished or
// wait for a reliable update freeze while the others iterate.
#include <type_traits>

#include "qed2d_cg.h"
#include "qed2d_fields.h"

namespace qed2d {

struct BWhere {
    int c;
    Latt L;
    long o;      // element offset of the chain's fields (2 * off)
    long uo;     // element offset of the chain's links (2 * uoff)
    int pb;      // first partial slot of the chain
    int nb;      // blocks of the chain in this launch
    double am0, beta;
    MarchPos p;  // p.blk: block index inside the chain
};

template <bool PAIR>
__device__ __forceinline__ BWhere bwhere(const BGeo &g, int R)
{
    const int b = xcd_remap(blockIdx.x, g.nblocks);  // (launches may add blocks past these)
    int lo = 0, local;
    if (g.inter) {
        lo = b % g.n;
        local = b / g.n;
    } else {
        int hi = g.n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (g.blk0[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        local = b - g.blk0[lo];
    }
    const BChain &ch = g.ch[lo];
    const int4 sh = g.shape[lo];
    BWhere w;
    w.c = lo;
    w.L = ch.L;
    w.o = 2 * ch.off;
    w.uo = 2 * ch.uoff;
    w.pb = ch.pb;
    w.nb = sh.w;
    w.am0 = ch.am0;
    w.beta = ch.beta;
    const int tx = threadIdx.x % sh.x, ty = threadIdx.x / sh.x;
    w.p.i1 = (local % sh.z) * sh.x + tx;
    w.p.r0 = ((local / sh.z) * sh.y + ty) * R;
    w.p.active = (w.p.i1 < (PAIR ? ch.L.L1 / 2 : ch.L.L1)) && (w.p.r0 < ch.L.L2);
    w.p.blk = local;
    return w;
}

// A chain takes a batched reliable update when its K1 paused for one
// (need_ru) or requested one and went on iterating (ru_req); neither is
// changed by the update's kernels (the K1 after it clears both), so every
// block of them decides alike.  Either way kk is the K1 that runs next, and
// K2(kk - 1) left the pending x update alpha_x p_{kk-1}.
__device__ __forceinline__ bool ru_wanted(const SolverState *S)
{
    return S->need_ru != 0 || S->ru_req != 0;
}

// Chain of arena element e (0 <= e < 2 * total sites).
__device__ __forceinline__ int chain_of_elem(const BChain *ch, int n, long e)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (2 * ch[mid].off <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// so = Q si on every chain (on the chains waiting for a reliable update only,
// when their solver states ru are given).
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(kBlock) kb_q_apply(BGeo g, const double2 *__restrict__ U,
                                                     const double2 *__restrict__ si,
                                                     double2 *__restrict__ so,
                                                     const SolverState *ru)
{
    const BWhere w = bwhere<false>(g, R);
    if (!w.p.active) return;
    if (ru && !ru[w.c].need_ru) return;
    SrcPlain<double> src{si + w.o, w.L.V};
    EpiStore<double> epi{so + w.o, w.L.V};
    q_march<double, R>(w.L, U + w.uo, 2.0 + w.am0, src, epi, w.p.i1, w.p.r0);
}
void launch_b_q_apply(const BGeo &g, int rows, const double2 *U, const double2 *si,
                      double2 *so, const SolverState *ru, hipStream_t st)
{
    DISPATCH_ROWS(rows, kb_q_apply, dim3(g.nblocks), dim3(kBlock), 0, st, g, U, si, so, ru);
}

// ---------------------------------------------------------------------------
// Gauge + fermion force and momentum update, one thread per arena site.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) kb_force_mom(const BChain *ch, int n, long vtot,
                                                       const double2 *__restrict__ U,
                                                       const double2 *__restrict__ X,
                                                       const double2 *__restrict__ Y,
                                                       double *__restrict__ mom, double eps)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < vtot;
         e += (long)gridDim.x * blockDim.x) {
        const int c = chain_of_elem(ch, n, 2 * e);
        const BChain &b = ch[c];
        const long s = e - b.off;
        const long o = 2 * b.off;
        force_mom_site(b.L, U + 2 * b.uoff, X ? X + o : nullptr, Y ? Y + o : nullptr, nullptr, mom + o,
                       b.beta, eps, 1.0, s);
    }
}
void launch_b_force_mom(const BChain *ch, int nchain, long vtot, const double2 *U,
                        const double2 *X, const double2 *Y, double *mom, double eps,
                        hipStream_t st)
{
    long nb = (vtot + kBlock - 1) / kBlock;
    if (nb > 8192) nb = 8192;
    launch_k(kb_force_mom, dim3((int)nb), dim3(kBlock), 0, st, ch, nchain, vtot, U, X,
                       Y, mom, eps);
}

// ---------------------------------------------------------------------------
// Momenta and the Nf=2 pseudofermion of every chain in one launch: each chain's
// k_fill_momenta and k_fill_cgauss (flavour 0) with the chain's own key
// (src/QED2dHMC.jl:69-70).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) kb_fill_mom_pf(const BChain *ch, int n, long vtot,
                                                         const uint2 *keys, uint32_t traj,
                                                         double *__restrict__ mom,
                                                         double2 *__restrict__ X)
{
#pragma clang fp contract(off)
    const double sq2 = 1.4142135623730951;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < vtot;
         e += (long)gridDim.x * blockDim.x) {
        const int c = chain_of_elem(ch, n, 2 * e);
        const BChain &b = ch[c];
        const long s = e - b.off, V = b.L.V, o = 2 * b.off;
        const uint2 k = keys[c];
        double g1, g2;
        gauss_pair((uint32_t)s, 0u, kPurposeMom, traj, k.x, k.y, &g1, &g2);
        mom[o + s] = g1;
        mom[o + V + s] = g2;
        for (uint32_t cc = 0; cc < 2; ++cc) {
            gauss_pair((uint32_t)s, cc, kPurposePF, traj, k.x, k.y, &g1, &g2);
            X[o + cc * V + s] = mk<double>(g1 / sq2, g2 / sq2);
        }
    }
}
void launch_b_fill_mom_pf(const BChain *ch, int nchain, long vtot, const uint2 *keys,
                          uint32_t traj, double *mom, double2 *X, hipStream_t st)
{
    long nb = (vtot + kBlock - 1) / kBlock;
    if (nb > 8192) nb = 8192;
    launch_k(kb_fill_mom_pf, dim3((int)nb), dim3(kBlock), 0, st, ch, nchain, vtot, keys, traj,
             mom, X);
}

// ---------------------------------------------------------------------------
// CG start from x = 0 (src/QED2dSolver.jl:10-16) on every chain.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) kb_cg_init(BGeo g, int rows,
                                                     const double2 *__restrict__ b,
                                                     double2 *__restrict__ x,
                                                     c2<T> *__restrict__ r,
                                                     c2<T> *__restrict__ p0,
                                                     float2 *__restrict__ xs, BParts P)
{
    const BWhere w = bwhere<false>(g, rows);
    const Latt &L = w.L;
    double acc = 0.0;
    if (w.p.active) {
        for (int kk = 0; kk < rows; ++kk) {
            const int row = w.p.r0 + kk;
            if (row >= L.L2) break;
            for (int c = 0; c < 2; ++c) {
                const long i = w.o + c * L.V + (long)row * L.L1 + w.p.i1;
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
        P.rr[0][w.pb + w.p.blk] = s;
        P.b[w.pb + w.p.blk] = s;
        if (w.p.blk == 0) {
            P.nrr[0][w.c] = w.nb;
            P.nb[w.c] = w.nb;
        }
    }
}
void launch_b_cg_init(const BGeo &g, int rows, const double2 *b, double2 *x, void *r, void *p0,
                      float2 *xs, int mixed, BParts P, hipStream_t st)
{
    if (mixed)
        launch_k(kb_cg_init<float>, dim3(g.nblocks), dim3(kBlock), 0, st, g, rows, b,
                           x, (float2 *)r, (float2 *)p0, xs, P);
    else
        launch_k(kb_cg_init<double>, dim3(g.nblocks), dim3(kBlock), 0, st, g, rows, b,
                           x, (double2 *)r, (double2 *)p0, (float2 *)nullptr, P);
}

// ---------------------------------------------------------------------------
// Residual from x with t = Q x: mode 0 starts a solve from a guess on every
// chain (FP64: r and p0 = 0; mixed: rs and p0s = 0); mode 2 is the reliable
// update of the chains that want one (ru_wanted: paused, or requested on the
// way).  Mode 2 resumes a paused chain (stop_k = 0) but leaves need_ru and
// ru_req to the K1 that runs next: a block of this kernel that starts after
// block 0 has finished must still see the request.
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(kBlock) kb_cg_resid(BGeo g, const double2 *__restrict__ U,
                                                      const double2 *__restrict__ t,
                                                      const double2 *__restrict__ b,
                                                      const double2 *__restrict__ x,
                                                      double2 *__restrict__ r,
                                                      float2 *__restrict__ rs,
                                                      double2 *__restrict__ p0,
                                                      float2 *__restrict__ p0s, SolverState *S,
                                                      BParts P, int mode, double2 *xout,
                                                      float2 *xs0)
{
    const BWhere w = bwhere<false>(g, R);
    SolverState *Sc = S + w.c;
    if (mode == 2 && !ru_wanted(Sc)) return;
    const int k = (mode == 2) ? Sc->kk : 1;
    const long o = w.o;
    EpiResid epi{b + o, x + o, r ? r + o : nullptr, rs ? rs + o : nullptr,
                 p0 ? p0 + o : nullptr, p0s ? p0s + o : nullptr, Sc->mu2, w.L.V, 0.0, 0.0,
                 xout ? xout + o : nullptr, xs0 ? xs0 + o : nullptr};
    if (w.p.active) {
        SrcPlain<double> src{t + o, w.L.V};
        q_march<double, R>(w.L, U + w.uo, 2.0 + w.am0, src, epi, w.p.i1, w.p.r0);
    }
    double a, bb;
    block_sum2_t0(epi.rr, epi.bb, a, bb);
    if (threadIdx.x == 0) {
        const int rd = (k - 1) & 1;
        P.rr[rd][w.pb + w.p.blk] = a;
        if (mode != 2) P.b[w.pb + w.p.blk] = bb;
        if (w.p.blk == 0) {
            P.nrr[rd][w.c] = w.nb;
            if (mode != 2) P.nb[w.c] = w.nb;
            Sc->true_at = k;
            if (mode == 2) {
                Sc->n_ru += 1;
                Sc->stop_k = 0;
            }
        }
    }
}
void launch_b_cg_resid(const BGeo &g, int rows, const double2 *U, const double2 *b,
                       const double2 *t, const double2 *x, double2 *r, float2 *rs, double2 *p0,
                       float2 *p0s, SolverState *S, BParts P, int mode, hipStream_t st,
                       double2 *xout, float2 *xs0)
{
    DISPATCH_ROWS(rows, kb_cg_resid, dim3(g.nblocks), dim3(kBlock), 0, st, g, U, t, b, x, r, rs,
                  p0, p0s, S, P, mode, xout, xs0);
}

// The first kernel of a batched reliable update, on the chains that wait for
// one (need_ru): t = Q x' and xn = x' with x' = x + double(xs + alpha_x p),
// the new FP64 solution, formed where it is read (with the pending update of
// the site-pair K1 that paused, in the expressions of the single-lattice
// k_ru_accum, so the bits are the same).  kb_cg_resid in mode 2 then takes x'
// from xn, stores it into x and clears xs: two launches where an accumulation
// pass, a Q apply and the residual took three (ensemble 925 -> 948
// chain-trajectories/s).
struct SrcXNew {
    const double2 *__restrict__ x;
    const float2 *__restrict__ xs;
    const float2 *__restrict__ p;
    float alpha;
    long V;
    struct Raw {
        double2 x0, x1;
        float2 s0, s1, p0, p1;
    };
    __device__ __forceinline__ Raw fetch(long i) const
    {
        return Raw{x[i], x[V + i], xs[i], xs[V + i], p[i], p[V + i]};
    }
    __device__ __forceinline__ void value(const Raw &w, double2 &a, double2 &b) const
    {
        const float2 d0 = xpa(w.s0, alpha, w.p0), d1 = xpa(w.s1, alpha, w.p1);
        a = mk<double>(w.x0.x + (double)d0.x, w.x0.y + (double)d0.y);
        b = mk<double>(w.x1.x + (double)d1.x, w.x1.y + (double)d1.y);
    }
};
struct EpiTX {
    double2 *__restrict__ t;
    double2 *__restrict__ xn;
    long V;
    __device__ __forceinline__ NoPre pre(long) const { return {}; }
    __device__ __forceinline__ void operator()(long s, NoPre, double2 o1, double2 o2, double2 c1,
                                               double2 c2v)
    {
        t[s] = o1;
        t[V + s] = o2;
        xn[s] = c1;
        xn[V + s] = c2v;
    }
};
template <int R>
__global__ void __launch_bounds__(kBlock) kb_ru_apply(BGeo g, const double2 *__restrict__ U,
                                                      const double2 *__restrict__ x,
                                                      const float2 *__restrict__ xs,
                                                      const float2 *__restrict__ P0,
                                                      const float2 *__restrict__ P1,
                                                      double2 *__restrict__ t,
                                                      double2 *__restrict__ xn,
                                                      const SolverState *S)
{
    const BWhere w = bwhere<false>(g, R);
    if (!w.p.active) return;
    const SolverState *Sc = S + w.c;
    if (!ru_wanted(Sc)) return;
    const long o = w.o;
    const float2 *pp = ((Sc->kk - 1) & 1) ? P1 : P0;
    SrcXNew src{x + o, xs + o, pp + o, (float)Sc->alpha_x, w.L.V};
    EpiTX epi{t + o, xn + o, w.L.V};
    q_march<double, R>(w.L, U + w.uo, 2.0 + w.am0, src, epi, w.p.i1, w.p.r0);
}
void launch_b_ru_apply(const BGeo &g, int rows, const double2 *U, const double2 *x,
                       const float2 *xs, const float2 *p0, const float2 *p1, double2 *t,
                       double2 *xn, const SolverState *S, hipStream_t st)
{
    DISPATCH_ROWS(rows, kb_ru_apply, dim3(g.nblocks), dim3(kBlock), 0, st, g, U, x, xs, p0, p1,
                  t, xn, S);
}

// ---------------------------------------------------------------------------
// CG iteration kernels, per-chain iteration index (see the file comment).
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) kb_cg1(BGeo g, const c2<T> *__restrict__ U,
                                                 const c2<T> *__restrict__ r, c2<T> *P0,
                                                 c2<T> *P1, c2<T> *__restrict__ t,
                                                 SolverState *S, BParts P)
{
    __shared__ double lds[kBlock / kWave + 1];
    const BWhere w = bwhere<false>(g, R);
    SolverState *Sc = S + w.c;
    if (Sc->stop_k != 0) return;
    const int k = Sc->kk;
    const int wp = k & 1, rd = (k - 1) & 1;
    if (w.p.blk == 0 && threadIdx.x == 0) Sc->k2 = k;
    double beta;
    if (!cg_k1_scalars<std::is_same<T, float>::value>(Sc, P.rr[rd] + w.pb, P.b + w.pb,
                                                         P.nrr[rd][w.c], P.nb[w.c], k, w.p.blk,
                                                         lds, beta))
        return;
    const long o = w.o;
    const c2<T> *pold = (rd ? P1 : P0) + o;
    c2<T> *pnew = (wp ? P1 : P0) + o;
    EpiK1<T> epi{pnew, t + o, w.L.V, 0.0, 0.0};
    if (w.p.active) {
        SrcXpay<T> src{r + o, pold, (T)beta, w.L.V};
        q_march<T, R>(w.L, U + w.uo, T(2) + (T)w.am0, src, epi, w.p.i1, w.p.r0);
    }
    double a, b;
    block_sum2_t0(epi.tt, epi.pp, a, b);
    if (threadIdx.x == 0) {
        P.tt[wp][2 * (w.pb + w.p.blk)] = a;
        P.tt[wp][2 * (w.pb + w.p.blk) + 1] = b;
        if (w.p.blk == 0) P.ntt[wp][w.c] = w.nb;
    }
}

template <typename T, int R>
__global__ void __launch_bounds__(kBlock) kb_cg2(BGeo g, const c2<T> *__restrict__ U,
                                                 const c2<T> *__restrict__ t, const c2<T> *P0,
                                                 const c2<T> *P1, c2<T> *__restrict__ x,
                                                 c2<T> *__restrict__ r, SolverState *S, BParts P)
{
    __shared__ double lds[kBlock / kWave + 1];
    const BWhere w = bwhere<false>(g, R);
    SolverState *Sc = S + w.c;
    if (Sc->stop_k != 0) return;
    const int k = Sc->k2;
    const int wp = k & 1;
    const double tt = sum_prev_partials(P.tt[wp] + 2 * w.pb, P.ntt[wp][w.c], 2, 0, lds);
    const double pp = sum_prev_partials(P.tt[wp] + 2 * w.pb, P.ntt[wp][w.c], 2, 1, lds);
    const double alpha = Sc->norm[wp] / (tt + Sc->mu2 * pp);
    const long o = w.o;
    EpiK2<T> epi{(wp ? P1 : P0) + o, x + o, r + o, (T)alpha, (T)Sc->mu2, w.L.V, 0.0};
    if (w.p.active) {
        SrcPlain<T> src{t + o, w.L.V};
        q_march<T, R>(w.L, U + w.uo, T(2) + (T)w.am0, src, epi, w.p.i1, w.p.r0);
    }
    const double a = block_sum_t0(epi.rr);
    if (threadIdx.x == 0) {
        P.rr[wp][w.pb + w.p.blk] = a;
        if (w.p.blk == 0) {
            P.nrr[wp][w.c] = w.nb;
            Sc->kk = k + 1;
        }
    }
}

template <int R, bool PK>
__global__ void __launch_bounds__(kBlock) kb_cg1v(BGeo g, LinksF32 U,
                                                  const float2 *__restrict__ r, float2 *P0,
                                                  float2 *P1, float2 *__restrict__ t,
                                                  float2 *__restrict__ x, SolverState *S,
                                                  BParts P, int soft)
{
    __shared__ double lds[kBlock / kWave + 1];
    const BWhere w = bwhere<true>(g, R);
    SolverState *Sc = S + w.c;
    if (Sc->stop_k != 0) return;
    const int k = Sc->kk;
    const int wp = k & 1, rd = (k - 1) & 1;
    if (w.p.blk == 0 && threadIdx.x == 0) Sc->k2 = k;
    double beta;
    bool fresh;
    if (!cg_k1_scalars<true>(Sc, P.rr[rd] + w.pb, P.b + w.pb, P.nrr[rd][w.c], P.nb[w.c], k,
                             w.p.blk, lds, beta, &fresh, soft != 0))
        return;
    const long o = w.o;
    const float2 *pold = (rd ? P1 : P0) + o;
    EpiK1v epi{(wp ? P1 : P0) + o, t + o, x + o, (float)Sc->alpha_x, !fresh, !fresh, w.L.V,
               0.0, 0.0};
    if (w.p.active) {
        SrcXpay2 src{r + o, pold, (float)beta, w.L.V};
        q_march2xt<R, PK>(w.L, U.at(w.uo), 2.0f + (float)w.am0, src, epi, w.p.i1, w.p.r0);
    }
    double a, b;
    block_sum2_t0(epi.tt, epi.pp, a, b);
    if (threadIdx.x == 0) {
        P.tt[wp][2 * (w.pb + w.p.blk)] = a;
        P.tt[wp][2 * (w.pb + w.p.blk) + 1] = b;
        if (w.p.blk == 0) P.ntt[wp][w.c] = w.nb;
    }
}

template <int R, bool PK, bool MU>
__device__ __forceinline__ void b_cg2v_body(const BGeo &g, const LinksF32 &U,
                                            const float2 *__restrict__ t, const float2 *P0,
                                            const float2 *P1, float2 *__restrict__ r,
                                            SolverState *S, const BParts &P)
{
    __shared__ double lds[kBlock / kWave + 1];
    const BWhere w = bwhere<true>(g, R);
    SolverState *Sc = S + w.c;
    // The chain's state goes out first through a per-lane address, the first
    // rows' loads next (they do not depend on the iteration); the partials of
    // K1 (slot by iteration) and alpha follow once the state is in.  A chain
    // that stopped returns before any store.  (With a shift, p_k is read with
    // the rows: its slot waits for the state.)
    const SolverState *Sv = state_v(Sc);
    const int skv = Sv->stop_k, kv = Sv->k2;
    const long o = w.o;
    EpiK2v<MU> epi{nullptr, r + o, 0.0f, 0.0f, w.L.V, 0.0};
    if (MU) epi.p = ((__builtin_amdgcn_readfirstlane(kv) & 1) ? P1 : P0) + o;
    int k = 0, wp = 0;
    double alpha = 0.0;
    auto ready = [&]() -> bool {
        if (__builtin_amdgcn_readfirstlane(skv) != 0) return false;
        k = __builtin_amdgcn_readfirstlane(kv);
        wp = k & 1;
        const double tt = sum_prev_partials(P.tt[wp] + 2 * w.pb, P.ntt[wp][w.c], 2, 0, lds);
        const double pp = sum_prev_partials(P.tt[wp] + 2 * w.pb, P.ntt[wp][w.c], 2, 1, lds);
        alpha = Sc->norm[wp] / (tt + Sc->mu2 * pp);
        epi.alpha = (float)alpha;
        epi.mu2 = (float)Sc->mu2;
        return true;
    };
    SrcPlain2 src{t + o, w.L.V};
    if (!q_march2t<R, PK>(w.L, U.at(w.uo), 2.0f + (float)w.am0, src, epi, w.p.i1, w.p.r0,
                          w.p.active, ready))
        return;
    const double a = block_sum_t0(epi.rr);
    if (threadIdx.x == 0) {
        P.rr[wp][w.pb + w.p.blk] = a;
        if (w.p.blk == 0) {
            P.nrr[wp][w.c] = w.nb;
            Sc->kk = k + 1;
            Sc->alpha_x = alpha;  // for K1(k + 1)
        }
    }
}
template <int R, bool PK>
__global__ void __launch_bounds__(kBlock) kb_cg2v(BGeo g, LinksF32 U,
                                                  const float2 *__restrict__ t, const float2 *P0,
                                                  const float2 *P1, float2 *__restrict__ r,
                                                  SolverState *S, BParts P)
{
    b_cg2v_body<R, PK, false>(g, U, t, P0, P1, r, S, P);
}
// (chains with a shift mu2: r -= alpha (Q t + mu2 p_k))
template <int R, bool PK>
__global__ void __launch_bounds__(kBlock) kb_cg2vm(BGeo g, LinksF32 U,
                                                   const float2 *__restrict__ t, const float2 *P0,
                                                   const float2 *P1, float2 *__restrict__ r,
                                                   SolverState *S, BParts P)
{
    b_cg2v_body<R, PK, true>(g, U, t, P0, P1, r, S, P);
}

void launch_b_cg_k1_f64(const BGeo &g, int rows, const double2 *U, const double2 *r,
                        double2 *p0, double2 *p1, double2 *t, SolverState *S, BParts P,
                        hipStream_t st)
{
    dim3 grid(g.nblocks), blk(kBlock);
    switch (rows) {
    case 1: launch_k((kb_cg1<double, 1>), grid, blk, 0, st, g, U, r, p0, p1, t, S, P); break;
    case 2: launch_k((kb_cg1<double, 2>), grid, blk, 0, st, g, U, r, p0, p1, t, S, P); break;
    case 4: launch_k((kb_cg1<double, 4>), grid, blk, 0, st, g, U, r, p0, p1, t, S, P); break;
    case 8: launch_k((kb_cg1<double, 8>), grid, blk, 0, st, g, U, r, p0, p1, t, S, P); break;
    default: launch_k((kb_cg1<double, 16>), grid, blk, 0, st, g, U, r, p0, p1, t, S, P); break;
    }
}
void launch_b_cg_k2_f64(const BGeo &g, int rows, const double2 *U, const double2 *t,
                        const double2 *p0, const double2 *p1, double2 *x, double2 *r,
                        SolverState *S, BParts P, hipStream_t st)
{
    dim3 grid(g.nblocks), blk(kBlock);
    switch (rows) {
    case 1: launch_k((kb_cg2<double, 1>), grid, blk, 0, st, g, U, t, p0, p1, x, r, S, P); break;
    case 2: launch_k((kb_cg2<double, 2>), grid, blk, 0, st, g, U, t, p0, p1, x, r, S, P); break;
    case 4: launch_k((kb_cg2<double, 4>), grid, blk, 0, st, g, U, t, p0, p1, x, r, S, P); break;
    case 8: launch_k((kb_cg2<double, 8>), grid, blk, 0, st, g, U, t, p0, p1, x, r, S, P); break;
    default: launch_k((kb_cg2<double, 16>), grid, blk, 0, st, g, U, t, p0, p1, x, r, S, P); break;
    }
}
void launch_b_cg_k1_f32(const BGeo &g, int rows, const void *U, bool packed, const float2 *r,
                        float2 *p0, float2 *p1, float2 *t, float2 *x, SolverState *S, BParts P,
                        hipStream_t st, bool soft)
{
    const LinksF32 lk{U, packed};
    DISPATCH_ROWS_PK(rows, packed, kb_cg1v, dim3(g.nblocks), dim3(kBlock), 0, st, g, lk, r, p0,
                     p1, t, x, S, P, soft ? 1 : 0);
}
// (x is updated by the next K1 or the reliable update's flush)
void launch_b_cg_k2_f32(const BGeo &g, int rows, const void *U, bool packed, const float2 *t,
                        const float2 *p0, const float2 *p1, float2 *, float2 *r,
                        SolverState *S, BParts P, bool mu, hipStream_t st)
{
    const LinksF32 lk{U, packed};
    if (mu) {
        DISPATCH_ROWS_PK(rows, packed, kb_cg2vm, dim3(g.nblocks), dim3(kBlock), 0, st, g, lk, t,
                         p0, p1, r, S, P);
    } else {
        DISPATCH_ROWS_PK(rows, packed, kb_cg2v, dim3(g.nblocks), dim3(kBlock), 0, st, g, lk, t,
                         p0, p1, r, S, P);
    }
}

// ---------------------------------------------------------------------------
// Multi-shift CG of several right-hand sides at once, one per chain (the
// pseudofermions of an RHMC force, sharing the links: uoff = 0).  Chain c's
// multi-shift state M[c] holds its shifts, its solution and direction
// buffers and its right-hand side; the iteration index is per chain as in
// the batched CG.  A chain that stopped at K1(s) has its K2(s) advance kk past
// s, so K1(s) never runs twice: the step in which the last shifts converge
// updates their solutions and must not repeat.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) kb_ms_init(BGeo g, int rows, c2<T> *__restrict__ r,
                                                     c2<T> *__restrict__ p0,
                                                     const MultiShiftState *M, BParts P)
{
    const BWhere w = bwhere<false>(g, rows);
    const MultiShiftState *Mc = M + w.c;
    const Latt &L = w.L;
    const int nshift = Mc->nshift;
    const c2<T> z = mk<T>(T(0), T(0));
    double acc = 0.0;
    if (w.p.active) {
        for (int kk = 0; kk < rows; ++kk) {
            const int row = w.p.r0 + kk;
            if (row >= L.L2) break;
            for (int c = 0; c < 2; ++c) {
                const long i = c * L.V + (long)row * L.L1 + w.p.i1;
                const double2 v = Mc->b[i];
                r[w.o + i] = from_d2<T>(v);
                p0[w.o + i] = z;
                for (int k = 0; k < nshift; ++k) {
                    ms_vec<T>(Mc->x[k])[i] = z;
                    ms_vec<T>(Mc->ps[k])[i] = z;
                }
                acc += v.x * v.x + v.y * v.y;
            }
        }
    }
    const double s = block_sum_t0(acc);
    if (threadIdx.x == 0) {
        P.rr[0][w.pb + w.p.blk] = s;
        if (w.p.blk == 0) P.nrr[0][w.c] = w.nb;
    }
}
void launch_b_ms_init(const BGeo &g, int rows, double2 *r, double2 *p0,
                      const MultiShiftState *M, BParts P, hipStream_t st)
{
    launch_k(kb_ms_init<double>, dim3(g.nblocks), dim3(kBlock), 0, st, g, rows, r, p0, M, P);
}
void launch_b_ms_init_f32(const BGeo &g, int rows, float2 *r, float2 *p0,
                          const MultiShiftState *M, BParts P, hipStream_t st)
{
    launch_k(kb_ms_init<float>, dim3(g.nblocks), dim3(kBlock), 0, st, g, rows, r, p0, M, P);
}

// Blocks past g.nblocks are shift-update blocks: chain c, shift s, elements
// eb * kShiftEPT * kBlock .. of the chain's spinor (maxshift * bpc of them per
// chain).
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) kb_ms1(BGeo g, const c2<T> *__restrict__ U,
                                                 const c2<T> *__restrict__ r, c2<T> *P0,
                                                 c2<T> *P1, c2<T> *__restrict__ t,
                                                 MultiShiftState *M, BParts P, int maxshift,
                                                 int bpc)
{
    if ((int)blockIdx.x >= g.nblocks) {
        const int idx = blockIdx.x - g.nblocks;
        const int c = idx / (maxshift * bpc), rem = idx - c * maxshift * bpc;
        const int sft = rem / bpc, eb = rem - sft * bpc;
        MultiShiftState *Mc = M + c;
        if (sft >= Mc->nshift) return;
        const int k = Mc->kk;
        const int sk = Mc->stop_k;
        if (sk != 0 && sk < k) return;
        const int rd = (k - 1) & 1;
        const BChain &ch = g.ch[c];
        ms1_shift_block<T>(Mc, r + 2 * ch.off, P.tt[rd] + 2 * ch.pb, P.rr[rd] + ch.pb,
                           P.nrr[rd][c], P.ntt[rd][c], k, Mc->nshift, sft,
                           (long)eb * kShiftEPT * kBlock, 2 * ch.L.V);
        return;
    }
    const BWhere w = bwhere<false>(g, R);
    MultiShiftState *Mc = M + w.c;
    const int k = Mc->kk;
    const int sk = Mc->stop_k;
    if (sk != 0 && sk < k) return;
    const int wp = k & 1, rd = (k - 1) & 1;
    if (w.p.blk == 0 && threadIdx.x == 0) {
        Mc->k2 = k;
        P.ntt[wp][w.c] = w.nb;
    }
    const long o = w.o;
    ms1_body<T, R>(w.L, U + w.uo, r + o, P0 + o, P1 + o, t + o, T(2) + (T)w.am0, Mc,
                   P.tt[wp] + 2 * w.pb, P.tt[rd] + 2 * w.pb, P.rr[rd] + w.pb, P.nrr[rd][w.c],
                   P.ntt[rd][w.c], k, Mc->nshift, w.p);
}

template <typename T, int R>
__global__ void __launch_bounds__(kBlock) kb_ms2(BGeo g, const c2<T> *__restrict__ U,
                                                 const c2<T> *__restrict__ t, const c2<T> *P0,
                                                 const c2<T> *P1, c2<T> *__restrict__ r,
                                                 MultiShiftState *M, BParts P)
{
    const BWhere w = bwhere<false>(g, R);
    MultiShiftState *Mc = M + w.c;
    const int k = Mc->k2;
    const int sk = Mc->stop_k;
    if (sk != 0 && sk <= k) {
        if (w.p.blk == 0 && threadIdx.x == 0) Mc->kk = k + 1;
        return;
    }
    const int wp = k & 1;
    const long o = w.o;
    ms2_body<T, R>(w.L, U + w.uo, t + o, P0 + o, P1 + o, r + o, T(2) + (T)w.am0, Mc,
                   P.tt[wp] + 2 * w.pb, P.rr[wp] + w.pb, P.ntt[wp][w.c], k, w.p);
    if (w.p.blk == 0 && threadIdx.x == 0) {
        P.nrr[wp][w.c] = w.nb;
        Mc->kk = k + 1;
    }
}
// DISPATCH_ROWS for the kernels templated on <spinor type, rows>.
#define DISPATCH_ROWS_T(rows, KERNEL, T, ...)                                         \
    switch (rows) {                                                                   \
    case 1: launch_k((KERNEL<T, 1>), __VA_ARGS__); break;                             \
    case 2: launch_k((KERNEL<T, 2>), __VA_ARGS__); break;                             \
    case 4: launch_k((KERNEL<T, 4>), __VA_ARGS__); break;                             \
    case 8: launch_k((KERNEL<T, 8>), __VA_ARGS__); break;                             \
    default: launch_k((KERNEL<T, 16>), __VA_ARGS__); break;                           \
    }

template <typename T>
static void b_ms_k1(const BGeo &g, int rows, const c2<T> *U, const c2<T> *r, c2<T> *p0,
                    c2<T> *p1, c2<T> *t, MultiShiftState *M, BParts P, long n2max,
                    int maxshift, hipStream_t st)
{
    const int bpc = (int)((n2max + (long)kShiftEPT * kBlock - 1) / ((long)kShiftEPT * kBlock));
    const long nb = g.nblocks + (long)g.n * maxshift * bpc;
    DISPATCH_ROWS_T(rows, kb_ms1, T, dim3((unsigned)nb), dim3(kBlock), 0, st, g, U, r, p0, p1, t,
                    M, P, maxshift, bpc);
}
void launch_b_ms_k1(const BGeo &g, int rows, const double2 *U, const double2 *r, double2 *p0,
                    double2 *p1, double2 *t, MultiShiftState *M, BParts P, long n2max,
                    int maxshift, hipStream_t st)
{
    b_ms_k1<double>(g, rows, U, r, p0, p1, t, M, P, n2max, maxshift, st);
}
void launch_b_ms_k1_f32(const BGeo &g, int rows, const float2 *U, const float2 *r, float2 *p0,
                        float2 *p1, float2 *t, MultiShiftState *M, BParts P, long n2max,
                        int maxshift, hipStream_t st)
{
    b_ms_k1<float>(g, rows, U, r, p0, p1, t, M, P, n2max, maxshift, st);
}
void launch_b_ms_k2(const BGeo &g, int rows, const double2 *U, const double2 *t,
                    const double2 *p0, const double2 *p1, double2 *r, MultiShiftState *M,
                    BParts P, hipStream_t st)
{
    DISPATCH_ROWS_T(rows, kb_ms2, double, dim3(g.nblocks), dim3(kBlock), 0, st, g, U, t, p0, p1,
                    r, M, P);
}
void launch_b_ms_k2_f32(const BGeo &g, int rows, const float2 *U, const float2 *t,
                        const float2 *p0, const float2 *p1, float2 *r, MultiShiftState *M,
                        BParts P, hipStream_t st)
{
    DISPATCH_ROWS_T(rows, kb_ms2, float, dim3(g.nblocks), dim3(kBlock), 0, st, g, U, t, p0, p1,
                    r, M, P);
}

}  // namespace qed2d
