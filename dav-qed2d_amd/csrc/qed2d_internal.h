// qed2d_internal.h -- host-side declarations shared between the kernel
// translation unit (qed2d_kernels.hip) and the orchestration / C-ABI layer.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/qed2d.h"
#include "qed2d_device.h"

namespace qed2d {

// FP32 links of the site-pair solver kernels: FP32 complex (float2 per link)
// or packed (one float per link, k_link_pack); at(o) is the field from link o.
struct LinksF32 {
    const void *U;
    int packed;
    __host__ __device__ LinksF32 at(long o) const
    {
        return LinksF32{packed ? (const void *)((const float *)U + o)
                               : (const void *)((const float2 *)U + o),
                        packed};
    }
};

// Kernel-class profiling (qed2d_profile_*): when armed, the next launch made
// through launch_k records these events at the start and at the end of the
// kernel's execution, not around its dispatch.  Host side, per thread.
struct ProfArm {
    hipEvent_t a, b;
};
extern thread_local ProfArm g_prof_arm;

// Every kernel of the library is launched through this.
template <typename K, typename... Args>
inline void launch_k(K kernel, const dim3 &grid, const dim3 &block, uint32_t shmem,
                     hipStream_t st, Args... args)
{
    if (g_prof_arm.a) {
        const ProfArm e = g_prof_arm;
        g_prof_arm = ProfArm{nullptr, nullptr};
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, e.a, e.b, 0u, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
    }
}

// Mixed-precision CG: a reliable update once the FP32 iterated residual has
// fallen by this factor (in norm) since the last one.  At 1e-3 the FP32
// residual still carries about four correct digits when it is replaced, the
// iteration counts equal FP64's, and the solutions differ from FP64's by far
// less than the stopping tolerance lets either move (4096^2 trajectory at tol
// 1e-20: dH 2e-8 from FP64's, against 1.2e-6 between FP64 at 1e-20 and 1e-24).
constexpr double kRuDelta = 1e-3;

// Device-resident solver state.  No kernel synchronises the grid: each
// kernel writes one partial sum per block, and the next kernel (after the
// kernel boundary) sums them in block order, so every block derives the same
// scalars.  Block 0 of the first CG kernel of iteration k commits the
// scalars later kernels need into slot k & 1, so that no kernel reads a slot
// that a block of the same kernel may be writing.
struct SolverState {
    double norm[2];    // |r_{k-1}|^2, committed by K1(k) into slot k & 1
    double rmax2[2];   // mixed precision: largest |r|^2 since the last reliable update
    double tol, bnorm; // committed by K1(1)
    double eps, mu2, delta2;  // set by the host
    double err;        // last |r|^2 seen by K1 (host information)
    int iter;          // iterations completed when K1 last committed
    int maxiter;
    int status;        // host information: 0 running, 1 converged, 2 iteration limit
    int need_ru;       // host information: mixed precision waits for a reliable update
    int n_ru;          // reliable updates performed
    int guess;         // the solve started from a caller's guess
    int true_at;       // K1(true_at) sees a true (FP64) residual norm
    int stop_k;        // nonzero: K1(stop_k) ended or paused the solve; later kernels return
    int kk, k2;        // batched solves: the iteration K1 runs next / K2 runs (per chain)
    // FP32 site-pair iterations: K2(k) leaves x += alpha_k p_k to K1(k+1), which
    // reads p_k anyway; alpha_x is that pending alpha_k (written by K2's block 0)
    double alpha_x;
    int bad_links;  // set by launch_link_pack: the packed FP32 links are not usable
    // On-device reliable updates (the mixed solver's K1 launched in soft mode):
    // a K1 whose iterated |r| has dropped by delta keeps iterating and sets
    // ru_req; the reliable-update kernels queued at a chunk start run iff
    // ru_req && !stop_k, and the K1 after them (a fresh true residual) clears
    // ru_req.  A drop 1e-4 below the trigger, a residual that looks converged
    // and the iteration limit pause the solve for the host instead.
    int ru_req;
    double r0norm;  // |r_0|^2, committed by K1(1) (host information)
    // FP32 site-pair iterations: alpha_j in slot j & 1 (written by K2(j)'s
    // block 0).  Only the K1 of an even k applies the pending x updates, those
    // of p_{k-2} (if K2(k-2) ran after the last true residual) and p_{k-1}, in
    // that order; the K1 of an odd k leaves its one pending update alone.
    double alpha_xk[2];
};

// A copy of the solver state for the host: *out = *S in pinned host memory,
// then v into *seq (release, system scope).  Taken by the K2 that ends a
// queued chunk (block 0, thread 0, after its last write to S); out == nullptr:
// no copy.
struct StateSnap {
    SolverState *out;
    int *seq;
    int v;
};

// Multi-shift CG state (all shifts sigma_i = mu_i^2 of (Q^2 + sigma_i) x_i = b on
// one Krylov sequence).  Slots as above: K1(k) handles Krylov index j = k - 1.
constexpr int kMaxShift = 32;
struct MultiShiftState {
    double norm[2];                  // |r_j|^2
    double alpha[2];                 // alpha_{j-1}
    double beta[2];                  // beta_j (p_j = r_j + beta_j p_{j-1})
    double zeta[2][kMaxShift];       // zeta^i_j
    double zeta_prev[2][kMaxShift];  // zeta^i_{j-1}
    int active[2][kMaxShift];        // shift i still iterating after step j
    double tol, bnorm, eps, sigma_b, err;
    double s[kMaxShift];             // sigma_i - sigma_b
    int iter, maxiter, status, nshift, stop_k, pad;
    int conv_iter[kMaxShift];
    // solution and search direction of shift i: double2 arrays, float2 in the
    // FP32 solves (ms_vec<T> gives the typed pointer)
    void *x[kMaxShift];
    void *ps[kMaxShift];
    const double2 *b;                // batched solves: the right-hand side
    int kk, k2;                      // batched solves: iteration K1 runs next / K2 runs
    // the shift updates of the step K1 last ran (written by its block 0, applied
    // by the shift-update kernel that follows it): x_i += sh_a[i] p_i and, where
    // sh_u[i] == 2, p_i = sh_z[i] r + sh_b[i] p_i; sh_u[i] == 0 skips shift i
    double sh_a[kMaxShift], sh_b[kMaxShift], sh_z[kMaxShift];
    int sh_u[kMaxShift];
};

// Per-block partial sums of the solver kernels.  tt: (|Q p|^2, |p|^2) per block,
// two slots by iteration parity; rr: |r|^2 per block, two slots; b: |b|^2.
// The n* fields record how many blocks the kernel that last filled a slot
// had (the FP32 pair kernels use a different grid from the FP64 ones); the
// launch functions keep them up to date in enqueue order.
struct CGPartials {
    double *tt[2];
    double *rr[2];
    double *b;
    int ntt[2], nrr[2], nb;
};

// Scalars of a reduction that the host reads back.
struct RedOut {
    double v[8];
};

enum { kBlock = 256 };

// Launch-shape description of the marching stencil kernels.
struct MarchGrid {
    int bx, by;          // threads of a block along i1 and along i2 (bx * by = kBlock)
    int nbx, nby, rows;  // blocks along i1, blocks along i2, rows marched per thread
    int nblocks() const { return nbx * nby; }
};
MarchGrid march_grid(const Latt &L, int rows);

// Coefficients of the fourth-order Omelyan-Mryglod-Folk integrator
// (src/QED2dHMC.jl:241-246; Omelyan et al., CPC 151 (2003) 272).
constexpr double OMF_R1 = 0.08398315262876693;
constexpr double OMF_R2 = 0.2539785108410595;
constexpr double OMF_R3 = 0.6822365335719091;
constexpr double OMF_R4 = -0.03230286765269967;

// H = sum p^2 / 2 + S_g of a context's lattice as a double-double (hi, lo).
int ctx_hamiltonian_dd(qed2d_ctx *c, const double *mom, const double *U, double *hi, double *lo);
// Re a^dag b (|a|^2 when b is null) of a context's lattice as a double-double.
int ctx_dot_dd(qed2d_ctx *c, const double *a, const double *b, double *hi, double *lo);
// queued forms: the reduction lands in *dst (pinned) when the stream gets there;
// after a synchronisation the *_from forms give what the calls above return
int ctx_queue_hamiltonian(qed2d_ctx *c, const double *mom, const double *U, RedOut *dst);
int ctx_queue_dot_dd(qed2d_ctx *c, const double *a, const double *b, RedOut *dst);
void ctx_hamiltonian_from(const qed2d_ctx *c, const RedOut &r, double *hi, double *lo);
void ctx_dot_dd_from(const RedOut &r, double *hi, double *lo);

// ---- batched ensembles ----------------------------------------------------
// A chain set: lattices whose fields are concatenated in one arena.  Chain
// c's spinor and momentum fields are reference-layout fields starting at
// element 2 * off (its two planes of V sites), its links at element 2 * uoff
// (uoff = off for independent chains; uoff = 0 for right-hand sides that share
// one gauge field); its solver state is S[c]; its partial sums occupy slots
// pb .. pb + (its blocks) - 1.
struct BChain {
    Latt L;
    long off, uoff;
    int pb;
    int pad;
    double am0, beta;
};
// Launch geometry of one kernel family over a chain set: chain c owns blocks
// blk0[c] .. blk0[c+1]-1 of the launch, tiled by shape[c] = (bx, by, nbx, nb).
struct BGeo {
    const BChain *ch;
    const int *blk0;
    const int4 *shape;
    int n;        // chains
    int nblocks;  // blocks of the launch (blk0[n])
    // nonzero for chains of one lattice shape: block b is local block b / n of
    // chain b % n, so the chains' blocks of one tile run side by side (and
    // right-hand sides sharing the links read them once into L2)
    int inter;
};
// Per-chain partial sums of the batched CG: values at pb + local block, and
// the producer's block count per chain (written by its local block 0).
struct BParts {
    double *tt[2], *rr[2], *b;
    int *ntt[2], *nrr[2], *nb;
};
// CG (Q^2 x_j = src_j from x = 0) on a right-hand-side set made by
// qed2d_batch_create_rhs, with the given tile rows and reliable-update factor;
// on failure *err receives the set's message.
int batch_rhs_cg(qed2d_batch *b, const double *U, double *x, const double *src, double am0,
                 int maxiter, double eps, int precision, int *iters, int rows_q, int rows,
                 int rows32, int rows_k2, int rows32_k2, double ru_delta, std::string *err);
// as batch_rhs_cg, with a mass and a shift (Q^2 + mu2[c], null: 0) per right-hand side
// (guess: each solve starts from the x it is given, else from 0)
int batch_rhs_cg_shifted(qed2d_batch *b, const double *U, double *x, const double *src,
                         const double *am0, const double *mu2, int maxiter, double eps,
                         int precision, int *iters, int rows_q, int rows, int rows32,
                         int rows_k2, int rows32_k2, double ru_delta, std::string *err,
                         bool guess = false);
// ru (may be null): the chains' solver states; apply only on the chains that
// wait for a reliable update (need_ru).
void launch_b_q_apply(const BGeo &g, int rows, const double2 *U, const double2 *si,
                      double2 *so, const SolverState *ru, hipStream_t st);
void launch_b_force_mom(const BChain *ch, int nchain, long vtot, const double2 *U,
                        const double2 *X, const double2 *Y, double *mom, double eps,
                        hipStream_t st);
void launch_b_cg_init(const BGeo &g, int rows, const double2 *b, double2 *x, void *r, void *p0,
                      float2 *xs, int mixed, BParts P, hipStream_t st);
void launch_b_cg_resid(const BGeo &g, int rows, const double2 *U, const double2 *b,
                       const double2 *t, const double2 *x, double2 *r, float2 *rs, double2 *p0,
                       float2 *p0s, SolverState *S, BParts P, int mode, hipStream_t st,
                       double2 *xout = nullptr, float2 *xs0 = nullptr);
// The first kernel of a batched reliable update (chains with need_ru):
// t = Q x', xn = x', x' = x + double(xs + alpha_x p_{stop_k - 1}).
void launch_b_ru_apply(const BGeo &g, int rows, const double2 *U, const double2 *x,
                       const float2 *xs, const float2 *p0, const float2 *p1, double2 *t,
                       double2 *xn, const SolverState *S, hipStream_t st);
void launch_b_cg_k1_f64(const BGeo &g, int rows, const double2 *U, const double2 *r,
                        double2 *p0, double2 *p1, double2 *t, SolverState *S, BParts P,
                        hipStream_t st);
void launch_b_cg_k2_f64(const BGeo &g, int rows, const double2 *U, const double2 *t,
                        const double2 *p0, const double2 *p1, double2 *x, double2 *r,
                        SolverState *S, BParts P, hipStream_t st);
// soft: a K1 whose iterated residual has dropped by delta requests a reliable
// update (ru_req) and the chain goes on (else it pauses, need_ru).
void launch_b_cg_k1_f32(const BGeo &g, int rows, const void *U, bool packed, const float2 *r,
                        float2 *p0, float2 *p1, float2 *t, float2 *x, SolverState *S, BParts P,
                        hipStream_t st, bool soft = false);
void launch_b_cg_k2_f32(const BGeo &g, int rows, const void *U, bool packed, const float2 *t,
                        const float2 *p0, const float2 *p1, float2 *x, float2 *r,
                        SolverState *S, BParts P, bool mu, hipStream_t st);
// Each chain's momenta and Nf=2 pseudofermion from its key keys[c] (one launch).
void launch_b_fill_mom_pf(const BChain *ch, int nchain, long vtot, const uint2 *keys,
                          uint32_t traj, double *mom, double2 *X, hipStream_t st);
// Per-chain double-double reductions in one launch (+ one finishing launch, one
// block per chain), bit-equal to each chain's single-lattice reduction: chain c
// owns producer blocks blk0[c] .. blk0[c+1]-1, blk0[c+1] - blk0[c] =
// chain_red_blocks(V_c) for the observables (H: out[c] as launch_observables)
// and chain_red_blocks(2 V_c) for dot products (out[c] as launch_dot_dd).
int chain_red_blocks(long n);
void launch_b_observables(const BChain *ch, const int *blk0, int n, int nblocks,
                          const double2 *U, const double *mom, double *part, RedOut *out,
                          hipStream_t st);
void launch_b_dot_dd(const BChain *ch, const int *blk0, int n, int nblocks, const double2 *a,
                     const double2 *b, double *part, RedOut *out, hipStream_t st);
// Multi-shift CG of one right-hand side per chain (M[c]: chain c's state).
void launch_b_ms_init(const BGeo &g, int rows, double2 *r, double2 *p0,
                      const MultiShiftState *M, BParts P, hipStream_t st);
void launch_b_ms_k1(const BGeo &g, int rows, const double2 *U, const double2 *r, double2 *p0,
                    double2 *p1, double2 *t, MultiShiftState *M, BParts P, long n2max,
                    int maxshift, hipStream_t st);
void launch_b_ms_k2(const BGeo &g, int rows, const double2 *U, const double2 *t,
                    const double2 *p0, const double2 *p1, double2 *r, MultiShiftState *M,
                    BParts P, hipStream_t st);
// the same in FP32 (spinors, links, x_i and p_i; scalars and sums in FP64)
void launch_b_ms_init_f32(const BGeo &g, int rows, float2 *r, float2 *p0,
                          const MultiShiftState *M, BParts P, hipStream_t st);
void launch_b_ms_k1_f32(const BGeo &g, int rows, const float2 *U, const float2 *r, float2 *p0,
                        float2 *p1, float2 *t, MultiShiftState *M, BParts P, long n2max,
                        int maxshift, hipStream_t st);
void launch_b_ms_k2_f32(const BGeo &g, int rows, const float2 *U, const float2 *t,
                        const float2 *p0, const float2 *p1, float2 *r, MultiShiftState *M,
                        BParts P, hipStream_t st);

// ---- launches (all asynchronous on `st`) ---------------------------------
void launch_q_apply_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                        double am0, int rows, hipStream_t st);
void launch_q_apply_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so,
                        float am0, int rows, hipStream_t st);
void launch_q_apply_mrhs_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                             int nrhs, double am0, int rows, hipStream_t st);
void launch_q_apply_mrhs_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so,
                             int nrhs, float am0, int rows, hipStream_t st);
// STREAM kernels: which 0 = copy c = a, 1 = triad c = a + b/2, 2 = read a
// (kStreamReadBlocks partial sums into part).
constexpr int kStreamReadBlocks = 1024;
void launch_stream(int which, const double2 *a, const double2 *b, double2 *c, double *part,
                   long n, hipStream_t st);
void launch_gamm5(const Latt &L, const double2 *si, double2 *so, hipStream_t st);
// so = Q si + mu^2 x  (x may alias nothing; used for Q^2 + mu^2 with si = Q x)
void launch_q_shift_f64(const Latt &L, const double2 *U, const double2 *si, const double2 *x,
                        double2 *so, double am0, double mu2, int rows, hipStream_t st);

// ---- CG on Q^2 + mu^2 (all take the marching grid of `rows`) -------------
// Start from x = 0: x = 0, r = b, p0 = 0; |b|^2 partials into rr[0] and b.
void launch_cg_init_f64(const Latt &L, const double2 *b, double2 *x, double2 *r, double2 *p0,
                        CGPartials &P, int rows, hipStream_t st);
// Mixed precision start from x = 0: x = 0, rs = float(b), xs = 0, p0 = 0.
void launch_cg_init_mixed(const Latt &L, const double2 *b, double2 *x, float2 *rs, float2 *xs,
                          float2 *p0, CGPartials &P, int rows, hipStream_t st);
// Residual from x with t = Q x already computed: r = b - (Q t + mu^2 x)
// (x may be null when mu^2 = 0: it is then not read).
// mode 0: FP64 start from a guess (writes r, zeroes p0, |b|^2 partials);
// mode 1: mixed start from a guess (writes rs, zeroes p0s, |b|^2 partials);
// mode 2: mixed reliable update before K1(k) (writes rs);
// mode 3: the on-device reliable update before K1(k) (writes rs while a soft
// request is pending; keeps the slot's block count and pads up to it).
// |r|^2 partials go to rr[(k - 1) & 1].
void launch_cg_resid(const Latt &L, const double2 *U, const double2 *b, const double2 *t,
                     const double2 *x, double2 *r, float2 *rs, double2 *p0, float2 *p0s,
                     double am0, SolverState *s, CGPartials &P, int k, int mode, int rows,
                     hipStream_t st);
// Iteration k: K1 = convergence test, p = r + beta p, t = Q p; K2 = x, r update.
void launch_cg_k1_f64(const Latt &L, const double2 *U, const double2 *r, double2 *p0,
                      double2 *p1, double2 *t, double am0, SolverState *s, CGPartials &P,
                      int k, int rows, hipStream_t st);
// (snap: the state copy of a chunk ending with this K2, see StateSnap)
void launch_cg_k2_f64(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                      const double2 *p1, double2 *x, double2 *r, double am0, SolverState *s,
                      CGPartials &P, int k, int rows, hipStream_t st,
                      StateSnap snap = StateSnap{nullptr, nullptr, 0});
// U: FP32 complex links, or (packed, site-pair kernels only: even L1 >= 4)
// links packed by launch_link_pack.
// (soft: on-device reliable updates, see SolverState::ru_req; guard, site-pair
// kernels only: K1(k) queued again behind a mode-4 reliable update, which it
// runs only if that update ran)
void launch_cg_k1_f32(const Latt &L, const void *U, bool packed, const float2 *r, float2 *p0,
                      float2 *p1, float2 *t, float2 *x, float am0, SolverState *s, CGPartials &P,
                      int k, int rows, hipStream_t st, bool soft = false, bool guard = false);
// (mu: the operator has mu2 != 0.  On even L1 >= 4 the site-pair kernels
// leave x += alpha p to the next K1, see SolverState::alpha_x; guard as K1's.)
void launch_cg_k2_f32(const Latt &L, const void *U, bool packed, const float2 *t,
                      const float2 *p0, const float2 *p1, float2 *x, float2 *r, float am0,
                      bool mu, SolverState *s, CGPartials &P, int k, int rows, hipStream_t st,
                      bool guard = false, StateSnap snap = StateSnap{nullptr, nullptr, 0});
// n FP64 U(1) links -> one float each (the packed format of the FP32 site-pair
// solver); sets S->bad_links if a link is not of unit modulus to 2^-23 (FP32 epsilon).
void launch_link_pack(const double2 *U, float *out, long n, SolverState *S, hipStream_t st);
// *out = *S in pinned host memory, then v into *seq (release, system scope)
void launch_state_snap(const SolverState *S, SolverState *out, int *seq, int v, hipStream_t st);
// A reliable update's x += double(xs), xs = 0; with p0 / p1 given (the FP32 site-pair solver, whose
// K1 was paused with an x update pending), xs + alpha_x p_{stop_k - 1} first.
// kb > 0: the on-device update queued before K1(kb), run only while a soft
// request is pending (pending update from p_{kb - 1}).  pk > 0: the update
// queued behind K1(pk), run only if K1(pk) paused (need_ru == pk).
void launch_ru_accum(const Latt &L, double2 *x, float2 *xs, const float2 *p0, const float2 *p1,
                     const SolverState *S, hipStream_t st, int kb = 0, int pk = 0);
// so = Q si while a soft reliable-update request is pending (SolverState::ru_req),
// or, pk > 0, while K1(pk) is paused for an update, or, pk < 0, once the solve
// has converged (status == 1).
void launch_q_apply_ru(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                       double am0, const SolverState *S, int rows, hipStream_t st, int pk = 0);
// Partials a consumer sums without a fold first (kPartMax of the device code).
constexpr int kPartsUnfolded = 1024;
// Block counts of the residual kernel and of the FP32 solver's K2 at a tile height.
int resid_blocks(const Latt &L, int rows);
int cg2_f32_blocks(const Latt &L, int rows);
void launch_to_f32(const double2 *in, float2 *out, long n, hipStream_t st);
void launch_to_f64(const float2 *in, double2 *out, long n, hipStream_t st);

// ---- multi-shift CG --------------------------------------------------------
// r = b, p0 = 0, x_i = 0, p_i = 0, |b|^2 partials into rr[0].
void launch_ms_init(const Latt &L, const double2 *b, double2 *r, double2 *p0, MultiShiftState *s,
                    CGPartials &P, int nshift, int rows, hipStream_t st);
void launch_ms_shifts(const MultiShiftState *s, const double2 *r, long n, int k, int nshift,
                      hipStream_t st);
void launch_ms_k1(const Latt &L, const double2 *U, const double2 *r, double2 *p0, double2 *p1,
                  double2 *t, double am0, MultiShiftState *s, CGPartials &P, int k,
                  int nshift, int rows, hipStream_t st);
void launch_ms_k2(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                  const double2 *p1, double2 *r, double am0, MultiShiftState *s,
                  CGPartials &P, int k, int nshift, int rows, hipStream_t st);

// Reductions (deterministic; results in out->v[]).
void launch_norm2(const double2 *a, long n, RedOut *out, double *part, hipStream_t st);
// a^dag b (|a|^2 when b is null) as double-doubles: out->v = (re hi, re lo,
// im hi, im lo); part holds 4 doubles per block (at most 1024 blocks).
void launch_dot_dd(const double2 *a, const double2 *b, long n, RedOut *out, double *part,
                   hipStream_t st);
void launch_dot(const double2 *a, const double2 *b, long n, RedOut *out, double *part,
                hipStream_t st);
// out->v[0..1] = sum Re P (double-double hi, lo); v[2..3] = sum arg P; v[4..5] = sum p^2
void launch_observables(const Latt &L, const double2 *U, const double *mom, double *plx,
                        double *qdens, RedOut *out, double *part, hipStream_t st);

// Forces and updates.
void launch_tr_dQwdU(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                     double *frc, hipStream_t st);
void launch_krnl_force(const Latt &L, const double2 *U, double beta, double *frc1, double *frc2,
                       hipStream_t st);
void launch_updt_fld(const Latt &L, double2 *U, const double *mom, double eps, hipStream_t st);
void launch_q_apply_tab(const Latt &L, const double2 *U, const double2 *const *X,
                        double2 *const *Y, const double *mass2, int n, int rows, hipStream_t st);
void launch_force_acc_tab(const Latt &L, const double2 *U, const double2 *const *X,
                          const double2 *const *Y, const double *rho, int n, double *facc,
                          hipStream_t st);
void launch_updt_gauge_mom(const Latt &L, const double2 *Ui, const double *mi, double2 *Uo,
                           double *mo, double eu, double beta, double ep, hipStream_t st);
// mom += eps * (Fgauge + cf * Ffermion(X, Y) + facc)   (facc may be null; X/Y null: gauge only)
void launch_force_mom(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                      const double *facc, double *mom, double beta, double eps, double cf,
                      hipStream_t st);
// facc (+)= rho * Ffermion(X, Y)
void launch_force_acc(const Latt &L, const double2 *U, const double2 *X, const double2 *Y,
                      double *facc, double rho, int first, hipStream_t st);

// Vector helpers.
void launch_axpy_c(double2 *y, const double2 *x, double ar, double ai, long n, hipStream_t st);
void launch_scale_add(double2 *y, double a, const double2 *x, double b, long n, hipStream_t st);
void launch_copy(double2 *dst, const double2 *src, long n, hipStream_t st);
// y = a y + b x, the coefficient `which` (0: a, 1: b) being 1 / red->v[0]
void launch_scale_add_inv(double2 *y, double a, const double2 *x, double b, const RedOut *red,
                          int which, long n, hipStream_t st);
void launch_negate(double *a, long n, hipStream_t st);

// RNG fills.
void launch_fill_momenta(const Latt &L, double *mom, uint32_t k0, uint32_t k1, uint32_t traj,
                         hipStream_t st);
void launch_fill_cgauss(const Latt &L, double2 *X, uint32_t k0, uint32_t k1, uint32_t traj,
                        uint32_t purpose, uint32_t sub0, hipStream_t st);
void launch_fill_pseudofermion(const Latt &L, double2 *X, uint32_t k0, uint32_t k1,
                               uint32_t traj, uint32_t flavour, hipStream_t st);
void launch_hot_start(const Latt &L, double2 *U, uint32_t k0, uint32_t k1, hipStream_t st);
void launch_unfold(const Latt &L, const double2 *Uin, double2 *Uout, int reflect,
                   hipStream_t st);

}  // namespace qed2d
