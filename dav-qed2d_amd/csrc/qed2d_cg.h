// qed2d_cg.h -- device pieces of the CG iteration shared by the single-lattice
// solvers (qed2d_solvers.hip) and the batched ensembles (qed2d_batch.hip):
// the convergence / reliable-update logic of K1 and the epilogues that fuse
// the vector algebra into the marching stencil (reference src/QED2dSolver.jl).
#pragma once

#include "qed2d_march.h"

namespace qed2d {

template <typename T> __device__ __forceinline__ c2<T> from_d2(double2 v)
{
    return mk<T>((T)v.x, (T)v.y);
}

// Convergence test and scalars at the start of K1(k), identical in every
// block: sums the |r_{k-1}|^2 partials, decides whether the solve ends
// (converged / iteration limit) or, for the FP32 iterations of the mixed
// solver, pauses for a reliable update; block 0 commits the scalars of slot
// k & 1.  Returns false when the kernel has no work; beta receives
// |r_{k-1}|^2 / |r_{k-2}|^2 (src/QED2dSolver.jl:33-35); fresh (if given) is
// set when r_{k-1} is a new start (k = 1 or right after a reliable update),
// i.e. when no deferred x update is pending.  cg_k1_decide takes
// |r_{k-1}|^2 (rr) already summed; cg_k1_scalars sums it first.
// The solver state K1(k) reads, loaded in one go (so a kernel can issue the
// loads early).
struct K1In {
    double bnorm, eps, tol, prev, rmax, delta2;
    int true_at, guess, maxiter;
};
__device__ __forceinline__ K1In k1_in(const SolverState *S, int k)
{
    const int rd = (k - 1) & 1;
    K1In v;
    v.bnorm = S->bnorm;
    v.eps = S->eps;
    v.tol = S->tol;
    v.prev = S->norm[rd];
    v.rmax = S->rmax2[rd];
    v.delta2 = S->delta2;
    v.true_at = S->true_at;
    v.guess = S->guess;
    v.maxiter = S->maxiter;
    return v;
}

// The same through vector loads: a kernel that issues them before its first
// rows' loads keeps them out of the scalar-load count, which the rows' address
// arithmetic (kernel arguments) waits on.
// The solver state seen through a per-lane address (lane offset 0, opaque to
// the compiler): its loads and values stay in vector registers, so nothing
// waits for them before their first use.
__device__ __forceinline__ const SolverState *state_v(const SolverState *S)
{
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return (const SolverState *)((const char *)S + z);
}

// The state copy for the host (StateSnap), by one thread after its last write
// to S: the copy, then the sequence number with a release at system scope, so
// that the host that reads the number finds the copy complete.
__device__ __forceinline__ void put_snap(const StateSnap &sn, const SolverState *S)
{
    if (sn.out == nullptr) return;
    *sn.out = *S;
    __hip_atomic_store(sn.seq, sn.v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// soft: the mixed solver's on-device reliable updates (SolverState::ru_req; a
// drop 1e-4 below the trigger pauses the solve instead).
template <bool mixed>
__device__ __forceinline__ bool cg_k1_decide(SolverState *S, const K1In &in, double rr,
                                             const double *Pb, int nbb, int k, int blk,
                                             double *lds, double &beta, bool *fresh = nullptr,
                                             bool soft = false)
{
    const int w = k & 1;
    const bool first = (k == 1);
    const double bn = first ? sum_prev_partials(Pb, nbb, 1, 0, lds) : in.bnorm;
    const double tol = first ? in.eps * bn : in.tol;
    const double prev = first ? 0.0 : in.prev;
    double rmax = first ? rr : in.rmax;
    const bool truer = first || in.true_at == k;
    if (fresh) *fresh = truer;
    int stat = 0, ru = 0, req = 0;
    if (first && bn == 0.0 && !in.guess) {
        stat = 1;  // b = 0: x = 0 exactly
    } else if (!mixed) {
        if (rr < tol) stat = 1;                         // src/QED2dSolver.jl:29
        else if (k - 1 >= in.maxiter) stat = 2;
    } else if (truer) {
        if (rr < tol) stat = 1;
        else {
            rmax = rr;
            if (k - 1 >= in.maxiter) stat = 2;
        }
    } else {
        if (rr < tol || k - 1 >= in.maxiter) ru = 1;
        else if (rr < in.delta2 * rmax) {
            // soft: keep iterating until a chunk start's reliable update, unless
            // the residual has fallen far past the trigger
            if (soft && !(rr < 1e-4 * in.delta2 * rmax)) req = 1;
            else ru = 1;
        } else if (rr > rmax) rmax = rr;
    }
    if (blk == 0 && threadIdx.x == 0) {
        S->err = rr;
        S->iter = k - 1;
        S->norm[w] = rr;
        S->rmax2[w] = rmax;
        if (first) {
            S->bnorm = bn;
            S->tol = tol;
            S->r0norm = rr;
        }
        if (soft) S->ru_req = req;  // (a fresh true residual clears it)
        if (stat) {
            S->status = stat;
            S->stop_k = k;
        }
        // need_ru: the iteration that paused (a batched or queued-ahead
        // reliable update leaves it set for the K1 it resumes to clear here)
        S->need_ru = ru ? k : 0;
        if (ru) S->stop_k = k;
    }
    if (stat || ru) return false;
    beta = first ? 0.0 : rr / prev;
    return true;
}

template <bool mixed>
__device__ __forceinline__ bool cg_k1_scalars(SolverState *S, const double *Prr, const double *Pb,
                                              int nrr, int nbb, int k, int blk, double *lds,
                                              double &beta, bool *fresh = nullptr,
                                              bool soft = false)
{
    const double rr = sum_prev_partials(Prr, nrr, 1, 0, lds);
    return cg_k1_decide<mixed>(S, k1_in(S, k), rr, Pb, nbb, k, blk, lds, beta, fresh, soft);
}

template <typename T> struct EpiK1 {
    c2<T> *__restrict__ pnew;
    c2<T> *__restrict__ t;
    long V;
    double tt, pp;
    __device__ __forceinline__ NoPre pre(long) const { return {}; }
    __device__ __forceinline__ void operator()(long s, NoPre, c2<T> o1, c2<T> o2, c2<T> c1,
                                               c2<T> cc2)
    {
        pnew[s] = c1;
        pnew[V + s] = cc2;
        t[s] = o1;
        t[V + s] = o2;
        tt += (double)(o1.x * o1.x + o1.y * o1.y) + (double)(o2.x * o2.x + o2.y * o2.y);
        pp += (double)(c1.x * c1.x + c1.y * c1.y) + (double)(cc2.x * cc2.x + cc2.y * cc2.y);
    }
};

template <typename T> struct EpiK2 {
    const c2<T> *__restrict__ p;
    c2<T> *__restrict__ x;
    c2<T> *__restrict__ r;
    T alpha, mu2;
    long V;
    double rr;
    struct Pre {
        c2<T> p1, p2, x1, x2, r1, r2;
    };
    __device__ __forceinline__ Pre pre(long s) const
    {
        return Pre{p[s], p[V + s], x[s], x[V + s], r[s], r[V + s]};
    }
    __device__ __forceinline__ void one(long i, c2<T> o, c2<T> pv, c2<T> xv, c2<T> rv)
    {
        const c2<T> a = mk<T>(o.x + mu2 * pv.x, o.y + mu2 * pv.y);
        xv = mk<T>(xv.x + alpha * pv.x, xv.y + alpha * pv.y);   // src/QED2dSolver.jl:24
        rv = mk<T>(rv.x - alpha * a.x, rv.y - alpha * a.y);     // src/QED2dSolver.jl:25
        x[i] = xv;
        r[i] = rv;
        rr += (double)(rv.x * rv.x + rv.y * rv.y);
    }
    __device__ __forceinline__ void operator()(long s, const Pre &q, c2<T> o1, c2<T> o2, c2<T>,
                                               c2<T>)
    {
        one(s, o1, q.p1, q.x1, q.r1);
        one(V + s, o2, q.p2, q.x2, q.r2);
    }
};

// x + alpha p of the FP32 site-pair solver: K1(k+1) and the reliable update's
// flush both apply the update that K2(k) deferred, with this one expression.
__device__ __forceinline__ float2 xpa(float2 x, float a, float2 p)
{
    return make_float2(__builtin_fmaf(a, p.x, x.x), __builtin_fmaf(a, p.y, x.y));
}

// K1(k) of the site-pair solver: stores p_k and t = Q p_k, sums |t|^2 and
// |p_k|^2, and unless r_{k-1} is fresh applies K2(k-1)'s deferred
// x += alpha_{k-1} p_{k-1} (src/QED2dSolver.jl:24) at the same sites.
struct EpiK1v {
    float2 *__restrict__ pnew;
    float2 *__restrict__ t;
    float2 *__restrict__ x;
    float alpha;  // alpha_{k-1}
    bool upd;     // apply the pending update (may be decided after the loads)
    bool ld;      // load x with the rows (set wherever upd may end up set)
    long V;
    double tt, pp;
    // Two pending updates (see SolverState::alpha_xk): p_{k-2} is read from
    // pnew at the thread's own sites before p_k overwrites it there (ldp2:
    // load it; two: apply alpha2 p_{k-2} first).
    float alpha2 = 0.0f;
    bool two = false;
    bool ldp2 = false;
    struct Pre {
        F2x2 x1, x2, p1, p2, q1, q2;
    };
    // p_{k-1} of the row comes with the source's raw loads (SrcXpay2::Raw2)
    template <typename Raw> __device__ __forceinline__ Pre pre_raw(long s, const Raw &w) const
    {
        if (!ld) return Pre{};
        Pre q{ld2(x, s), ld2(x, V + s), w.p0, w.p1, F2x2{}, F2x2{}};
        if (ldp2) {
            q.q1 = ld2(pnew, s);
            q.q2 = ld2(pnew, V + s);
        }
        return q;
    }
    __device__ __forceinline__ void xupd(long i, F2x2 xv, F2x2 pv, F2x2 qv)
    {
        if (two) {
            xv.a0 = xpa(xv.a0, alpha2, qv.a0);
            xv.a1 = xpa(xv.a1, alpha2, qv.a1);
        }
        st2(x, i, xpa(xv.a0, alpha, pv.a0), xpa(xv.a1, alpha, pv.a1));
    }
    __device__ __forceinline__ void pair(long s, const Pre &q, float2 oa1, float2 oa2, float2 ob1,
                                         float2 ob2, F2x2 c1, F2x2 c2)
    {
        st2(pnew, s, c1.a0, c1.a1);
        st2(pnew, V + s, c2.a0, c2.a1);
        st2(t, s, oa1, ob1);
        st2(t, V + s, oa2, ob2);
        if (upd) {
            xupd(s, q.x1, q.p1, q.q1);
            xupd(V + s, q.x2, q.p2, q.q2);
        }
        tt += (double)(cabs2<float>(oa1) + cabs2<float>(ob1)) +
              (double)(cabs2<float>(oa2) + cabs2<float>(ob2));
        pp += (double)(cabs2<float>(c1.a0) + cabs2<float>(c1.a1)) +
              (double)(cabs2<float>(c2.a0) + cabs2<float>(c2.a1));
    }
};

// K2(k) of the site-pair solver: r -= alpha_k (Q t + mu2 p_k)
// (src/QED2dSolver.jl:25); the x update is left to K1(k+1).  MU: mu2 != 0
// (only then is p_k read here).
template <bool MU> struct EpiK2v {
    const float2 *__restrict__ p;
    float2 *__restrict__ r;
    float alpha, mu2;
    long V;
    double rr;
    struct Pre {
        F2x2 p1, p2, r1, r2;
    };
    __device__ __forceinline__ Pre pre(long s) const
    {
        Pre q{};
        if (MU) {
            q.p1 = ld2(p, s);
            q.p2 = ld2(p, V + s);
        }
        q.r1 = ld2(r, s);
        q.r2 = ld2(r, V + s);
        return q;
    }
    __device__ __forceinline__ void comp(long i, float2 oa, float2 ob, F2x2 pv, F2x2 rv)
    {
        if (MU) {
            oa = make_float2(oa.x + mu2 * pv.a0.x, oa.y + mu2 * pv.a0.y);
            ob = make_float2(ob.x + mu2 * pv.a1.x, ob.y + mu2 * pv.a1.y);
        }
        rv.a0 = make_float2(rv.a0.x - alpha * oa.x, rv.a0.y - alpha * oa.y);
        rv.a1 = make_float2(rv.a1.x - alpha * ob.x, rv.a1.y - alpha * ob.y);
        st2(r, i, rv.a0, rv.a1);
        rr += (double)(cabs2<float>(rv.a0) + cabs2<float>(rv.a1));
    }
    __device__ __forceinline__ void pair(long s, const Pre &q, float2 oa1, float2 oa2, float2 ob1,
                                         float2 ob2, F2x2, F2x2)
    {
        comp(s, oa1, ob1, q.p1, q.r1);
        comp(V + s, oa2, ob2, q.p2, q.r2);
    }
};

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
    // (batched reliable update: x read from the updated copy is stored back
    // into xout, and the FP32 partial solution xs0 cleared; null otherwise)
    double2 *__restrict__ xout = nullptr;
    float2 *__restrict__ xs0 = nullptr;
    struct Pre {
        double2 b1, b2, x1, x2;
    };
    // (x == nullptr: x = 0, for mu2 == 0, so that x is not read)
    __device__ __forceinline__ Pre pre(long s) const
    {
        if (x == nullptr) return Pre{b[s], b[V + s], double2{}, double2{}};
        return Pre{b[s], b[V + s], x[s], x[V + s]};
    }
    __device__ __forceinline__ void one(long i, double2 o, double2 bv, double2 xv)
    {
        const double2 rv = mk<double>(bv.x - (o.x + mu2 * xv.x), bv.y - (o.y + mu2 * xv.y));
        if (r) r[i] = rv;
        if (rs) rs[i] = mk<float>((float)rv.x, (float)rv.y);
        if (p0) p0[i] = mk<double>(0.0, 0.0);
        if (p0s) p0s[i] = mk<float>(0.0f, 0.0f);
        if (xout) xout[i] = xv;
        if (xs0) xs0[i] = mk<float>(0.0f, 0.0f);
        rr += rv.x * rv.x + rv.y * rv.y;
        bb += bv.x * bv.x + bv.y * bv.y;
    }
    __device__ __forceinline__ void operator()(long s, const Pre &q, double2 o1, double2 o2,
                                               double2, double2)
    {
        one(s, o1, q.b1, q.x1);
        one(V + s, o2, q.b2, q.x2);
    }
};

// ---------------------------------------------------------------------------
// Multi-shift CG iteration bodies (the kernels k_ms1 / k_ms2 of one lattice
// and the batched kernels call them with the lattice, pointers and position
// of the calling block).
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ c2<T> *ms_vec(void *p)
{
    return static_cast<c2<T> *>(p);
}

// K1's epilogue: stores p_k and t = Q p_k, sums |t|^2 and |p_k|^2 (FP32
// solves: each site's sum in FP32, accumulated in FP64).
template <typename T> struct EpiMS1 {
    c2<T> *__restrict__ pnew;
    c2<T> *__restrict__ t;
    long V;
    double tt, pp;
    __device__ __forceinline__ NoPre pre(long) const { return {}; }
    __device__ __forceinline__ void operator()(long s, NoPre, c2<T> o1, c2<T> o2, c2<T> c1,
                                               c2<T> cc2)
    {
        pnew[s] = c1;
        pnew[V + s] = cc2;
        t[s] = o1;
        t[V + s] = o2;
        tt += (double)((o1.x * o1.x + o1.y * o1.y) + (o2.x * o2.x + o2.y * o2.y));
        pp += (double)((c1.x * c1.x + c1.y * c1.y) + (cc2.x * cc2.x + cc2.y * cc2.y));
    }
};

// The shift updates of one Krylov step at spinor element i (after K1 of that
// step has committed the coefficients): x_i += alpha^i p^i and, for shifts
// still iterating, p^i = zeta^i r + beta^i p^i.  Shift k only; elementwise and
// in place, so any number of elements and shifts run in parallel.
__device__ __forceinline__ void ms_shift_update(const MultiShiftState *S,
                                                const double2 *__restrict__ r, long i, int k)
{
    const int u = S->sh_u[k];
    if (u == 0) return;
    double2 *__restrict__ ps = ms_vec<double>(S->ps[k]);
    double2 *__restrict__ x = ms_vec<double>(S->x[k]);
    const double a = S->sh_a[k];
    const double2 pv = ps[i];
    const double2 xv = x[i];
    x[i] = mk<double>(xv.x + a * pv.x, xv.y + a * pv.y);
    if (u == 2) {
        const double z = S->sh_z[k], b = S->sh_b[k];
        const double2 rv = r[i];
        ps[i] = mk<double>(z * rv.x + b * pv.x, z * rv.y + b * pv.y);
    }
}
// Block-shared scalars of multi-shift K1(k).
struct MS1Shared {
    double lds[kBlock / kWave + 1];
    double zj[kMaxShift], z1[kMaxShift], as[kMaxShift], bs[kMaxShift];
    int upd[kMaxShift];
    int anyleft;
};

// The scalars of K1(k) (Krylov index j = k - 1), the same in every block:
// the base system's alpha_{j-1} and beta_j from the partials of the previous
// kernels, and per shift zeta_j, alpha^i, beta^i and its update (0: stopped
// earlier, 1: converges at this step, only x updates, 2: x and p update).
// Returns the status: 0 running, 1 every shift converged, 2 iteration limit.
__device__ __forceinline__ int ms1_scalars(const MultiShiftState *S, const double *Ptt_rd,
                                           const double *Prr_rd, int nrr, int ntt, int k,
                                           int nshift, MS1Shared &sh, double &rr, double &tol,
                                           double &alpha_jm1, double &beta_j)
{
    const int rd = (k - 1) & 1;
    rr = sum_prev_partials(Prr_rd, nrr, 1, 0, sh.lds);
    const bool first = (k == 1);
    if (first) {
        tol = S->eps * rr;
        alpha_jm1 = 1.0;
        beta_j = 0.0;
    } else {
        const double tt = sum_prev_partials(Ptt_rd, ntt, 2, 0, sh.lds);
        const double pp = sum_prev_partials(Ptt_rd, ntt, 2, 1, sh.lds);
        tol = S->tol;
        alpha_jm1 = S->norm[rd] / (tt + S->sigma_b * pp);
        beta_j = rr / S->norm[rd];
    }
    if (threadIdx.x < nshift) {
        const int i = threadIdx.x;
        double zj = 1.0, z1 = 1.0, as = 0.0, bs = 0.0;
        int act = 1, conv = 0;
        if (first) {
            conv = (rr == 0.0);
        } else {
            act = S->active[rd][i];
            if (act) {
                z1 = S->zeta[rd][i];
                const double z2 = S->zeta_prev[rd][i];
                const double alpha_jm2 = S->alpha[rd], beta_jm1 = S->beta[rd];
                zj = z1 * z2 * alpha_jm2 /
                     (alpha_jm1 * beta_jm1 * (z2 - z1) + z2 * alpha_jm2 * (1.0 + S->s[i] * alpha_jm1));
                as = alpha_jm1 * zj / z1;
                bs = beta_j * (zj / z1) * (zj / z1);
                conv = zj * zj * rr < tol;
            }
        }
        sh.zj[i] = zj;
        sh.z1[i] = z1;
        sh.as[i] = as;
        sh.bs[i] = bs;
        sh.upd[i] = act ? (conv ? 1 : 2) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int any = 0;
        for (int i = 0; i < nshift; ++i) any |= (sh.upd[i] == 2);
        sh.anyleft = any;
    }
    __syncthreads();
    return sh.anyleft ? ((k - 1 >= S->maxiter) ? 2 : 0) : 1;
}

template <typename T, int R>
__device__ __forceinline__ void ms1_body(const Latt &L, const c2<T> *__restrict__ U,
                                         const c2<T> *__restrict__ r, c2<T> *P0, c2<T> *P1,
                                         c2<T> *__restrict__ t, T mass2, MultiShiftState *S,
                                         double *Ptt_w, const double *Ptt_rd,
                                         const double *Prr_rd, int nrr, int ntt, int k,
                                         int nshift, const MarchPos &p)
{
    __shared__ MS1Shared sh;
    const int w = k & 1, rd = (k - 1) & 1;
    const bool first = (k == 1);
    double rr, tol, alpha_jm1, beta_j;
    const int stat = ms1_scalars(S, Ptt_rd, Prr_rd, nrr, ntt, k, nshift, sh, rr, tol, alpha_jm1,
                                 beta_j);
    if (p.blk == 0 && threadIdx.x == 0) {
        S->norm[w] = rr;
        S->alpha[w] = alpha_jm1;
        S->beta[w] = beta_j;
        for (int i = 0; i < nshift; ++i) {
            S->zeta[w][i] = sh.zj[i];
            S->zeta_prev[w][i] = sh.z1[i];
            S->active[w][i] = (sh.upd[i] == 2);
            if (sh.upd[i] == 1) S->conv_iter[i] = k - 1;
            // for the shift-update kernel that follows (none at the iteration limit)
            S->sh_u[i] = (stat == 2) ? 0 : sh.upd[i];
            S->sh_a[i] = sh.as[i];
            S->sh_b[i] = sh.bs[i];
            S->sh_z[i] = sh.zj[i];
        }
        S->iter = k - 1;
        S->err = rr;
        if (first) {
            S->tol = tol;
            S->bnorm = rr;
        }
        if (stat) {
            S->status = stat;
            S->stop_k = k;
        }
    }
    if (stat == 2 || !sh.anyleft) return;  // (the last x updates are the shift updates')
    const c2<T> *pold = rd ? P1 : P0;
    c2<T> *pnew = w ? P1 : P0;
    EpiMS1<T> epi{pnew, t, L.V, 0.0, 0.0};
    if (p.active) {
        SrcXpay<T> src{r, pold, (T)beta_j, L.V};
        q_march<T, R>(L, U, mass2, src, epi, p.i1, p.r0);
    }
    double a, b;
    block_sum2_t0(epi.tt, epi.pp, a, b);
    if (threadIdx.x == 0) {
        Ptt_w[2 * p.blk] = a;
        Ptt_w[2 * p.blk + 1] = b;
    }
}

// Spinor elements per thread of a shift-update block.
constexpr int kShiftEPT = 8;

// The shift updates of K1(k) run by extra blocks of the same launch (the
// batched K1): each derives the step's coefficients with ms1_scalars, as the
// stencil blocks do, and updates shift `shift` at spinor elements
// e0 .. e0 + kShiftEPT * blockDim - 1 (of n2): x_i += alpha^i p_i, and
// p_i = zeta r + beta^i p_i while the shift iterates (the expressions of
// ms_shift_update, with the coefficients rounded to T).  K1 writes none of
// r, x_i, p_i, so these blocks may run beside its stencil blocks.
template <typename T>
__device__ __forceinline__ void ms1_shift_block(const MultiShiftState *S,
                                                const c2<T> *__restrict__ r,
                                                const double *Ptt_rd, const double *Prr_rd,
                                                int nrr, int ntt, int k, int nshift, int shift,
                                                long e0, long n2)
{
    __shared__ MS1Shared sh;
    double rr, tol, alpha_jm1, beta_j;
    const int stat = ms1_scalars(S, Ptt_rd, Prr_rd, nrr, ntt, k, nshift, sh, rr, tol, alpha_jm1,
                                 beta_j);
    const int u = sh.upd[shift];
    if (stat == 2 || u == 0) return;
    c2<T> *__restrict__ ps = ms_vec<T>(S->ps[shift]);
    c2<T> *__restrict__ x = ms_vec<T>(S->x[shift]);
    const T a = (T)sh.as[shift], z = (T)sh.zj[shift], b = (T)sh.bs[shift];
#pragma unroll 4
    for (int j = 0; j < kShiftEPT; ++j) {
        const long i = e0 + (long)j * kBlock + threadIdx.x;
        if (i >= n2) break;
        const c2<T> pv = ps[i];
        const c2<T> xv = x[i];
        x[i] = mk<T>(xv.x + a * pv.x, xv.y + a * pv.y);
        if (u == 2) {
            const c2<T> rv = r[i];
            ps[i] = mk<T>(z * rv.x + b * pv.x, z * rv.y + b * pv.y);
        }
    }
}

// K2's epilogue: r -= alpha (Q t + sigma_b p), |r|^2 (FP32 solves: each
// element's square in FP32, accumulated in FP64).
template <typename T> struct EpiMS2 {
    c2<T> *__restrict__ r;
    const c2<T> *__restrict__ p;
    T alpha, sigma;
    long V;
    double rr;
    struct Pre {
        c2<T> p1, p2, r1, r2;
    };
    __device__ __forceinline__ Pre pre(long s) const { return Pre{p[s], p[V + s], r[s], r[V + s]}; }
    __device__ __forceinline__ void one(long i, c2<T> o, c2<T> pv, c2<T> rv)
    {
        rv = mk<T>(rv.x - alpha * (o.x + sigma * pv.x), rv.y - alpha * (o.y + sigma * pv.y));
        r[i] = rv;
        rr += (double)(rv.x * rv.x + rv.y * rv.y);
    }
    __device__ __forceinline__ void operator()(long s, const Pre &q, c2<T> o1, c2<T> o2, c2<T>,
                                               c2<T>)
    {
        one(s, o1, q.p1, q.r1);
        one(V + s, o2, q.p2, q.r2);
    }
};

template <typename T, int R>
__device__ __forceinline__ void ms2_body(const Latt &L, const c2<T> *__restrict__ U,
                                         const c2<T> *__restrict__ t, const c2<T> *P0,
                                         const c2<T> *P1, c2<T> *__restrict__ r, T mass2,
                                         MultiShiftState *S, const double *Ptt_w, double *Prr_w,
                                         int ntt, int k, const MarchPos &p)
{
    __shared__ double lds[kBlock / kWave + 1];
    const int w = k & 1;
    const double nw = S->norm[w], sb = S->sigma_b;
    // The partials' loads go out first (when they fit the per-lane buffer), the
    // first rows' loads next; alpha is formed once both are in flight, with
    // the same sums in the same order either way.
    const bool ahead = ntt <= kPartMax;
    PartialsAhead2 ptt;
    if (ahead) ptt.load(Ptt_w, ntt);
    EpiMS2<T> epi{r, w ? P1 : P0, T(0), (T)sb, L.V, 0.0};
    auto ready = [&]() -> bool {
        double tt, pp;
        if (ahead) {
            ptt.sum(tt, pp);
        } else {
            tt = sum_prev_partials(Ptt_w, ntt, 2, 0, lds);
            pp = sum_prev_partials(Ptt_w, ntt, 2, 1, lds);
        }
        epi.alpha = (T)(nw / (tt + sb * pp));
        return true;
    };
    SrcPlain<T> src{t, L.V};
    q_march_r<T, R>(L, U, mass2, src, epi, p.i1, p.r0, p.active, ready);
    const double a = block_sum_t0(epi.rr);
    if (threadIdx.x == 0) Prr_w[p.blk] = a;
}

}  // namespace qed2d
