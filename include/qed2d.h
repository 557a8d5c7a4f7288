/*
 * qed2d.h -- C ABI of libqed2d_hip, the MI355X (gfx950) implementation of the
 * QED2d molecular-dynamics hot path of dalbandea/Dav-QED2d (Julia module
 * QED2d).  Each entry point names the reference function it replaces.
 *
 * Fields
 *   All field pointers are DEVICE pointers in the reference's own layout,
 *   the Julia column-major arrays:
 *     spinor, link   ComplexF64[L1,L2,2]  -> double[2*2*V], index
 *                    2*((c*L2 + i2)*L1 + i1) + {0 re, 1 im}
 *     momentum/force Float64[L1,L2,2]     -> double[2*V]
 *     site density   Float64[L1,L2]       -> double[V]
 *   A CuArray/HIP buffer of the reference can therefore be passed as is.
 *
 * Contexts
 *   A context binds one lattice (L1, L2, beta, fermion boundary flags) to one
 *   HIP stream on the current device and owns every scratch buffer; it is the
 *   counterpart of the reference's LattParm (src/QED2d.jl:5-8).  The
 *   reference's KernelParm (src/QED2d.jl:10-13) has no counterpart: launch
 *   shapes are chosen by the library.  A context is not re-entrant; use one
 *   context per host thread.
 *
 * Calls are synchronous by default (results visible when the call returns),
 * as the reference's CUDA.@sync launches.  qed2d_ctx_set_sync(ctx, 0) makes
 * pure kernel launches asynchronous on the context's stream.
 *
 * Errors: every function returns QED2D_OK (0) or a negative status;
 * qed2d_last_error(ctx) gives the message (qed2d_last_error(NULL): why the
 * last qed2d_ctx_create of the calling thread failed).  The reference throws
 * (src/QED2dSolver.jl:40-43,99-102; src/QED2dRHMC.jl:12-14).
 *
 * Aliasing: an output may not alias an input of the same call
 * (QED2D_EINVAL), except where stated.  The reference's only aliased call
 * (src/QED2dRHMC.jl:159) is a data race there.
 */
#ifndef QED2D_H
#define QED2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QED2D_OK 0
#define QED2D_EINVAL (-1)   /* bad argument (sizes, aliasing, null pointers) */
#define QED2D_EHIP (-2)     /* HIP runtime error */
#define QED2D_ENOCONV (-3)  /* solver hit maxiter without converging */
#define QED2D_ENOMEM (-4)   /* device allocation failed */

/* Fermion boundary flag: antiperiodic across i2 = L2 -> 1 (direction 2,
 * "time").  The reference is periodic in both directions
 * (src/QED2dDirac.jl:7-11); 0 reproduces it. */
#define QED2D_BC_APERIODIC_T 1

/* Solver operator selectors: the function the reference passes as `A`. */
#define QED2D_OP_SQR_MSQ 0   /* gamm5Dw_sqr_msq : Q^2          (src/QED2dDirac.jl:69-81) */
#define QED2D_OP_SQR_MUSQ 1  /* gamm5Dw_sqr_musq: Q^2 + mu^2   (src/QED2dDirac.jl:83-95) */

/* Solver precisions. */
#define QED2D_PREC_F64 0     /* FP64 throughout (the reference) */
#define QED2D_PREC_MIXED 1   /* FP32 iterations + FP64 reliable updates, same FP64 stopping rule */

/* Integrators (src/integrators/integrators.jl:1-3). */
#define QED2D_LEAPFROG 0
#define QED2D_OMF4 1

typedef struct qed2d_ctx qed2d_ctx;

/* ---- context ------------------------------------------------------------
 * LattParm(iL, beta) (src/QED2d.jl:5-8) + the fermion BC flag. `stream` is a
 * hipStream_t (NULL = default stream). */
int qed2d_ctx_create(qed2d_ctx **ctx, int L1, int L2, double beta, int bc_flags, void *stream);
void qed2d_ctx_destroy(qed2d_ctx *ctx);
const char *qed2d_last_error(const qed2d_ctx *ctx);
int qed2d_ctx_set_beta(qed2d_ctx *ctx, double beta);
int qed2d_ctx_set_sync(qed2d_ctx *ctx, int sync);
/* Rows marched per thread by the FP64 / the FP32 stencil kernels (1, 2, 4, 8
 * or 16). */
int qed2d_ctx_set_tile(qed2d_ctx *ctx, int rows);
int qed2d_ctx_set_tile_f32(qed2d_ctx *ctx, int rows);
/* Tile heights per kernel family (rows marched per thread: 1, 2, 4, 8 or 16;
 * 0 leaves a family unchanged): FP64 Q/Dw applies, the FP32 Q apply, the FP64
 * solver kernels and the FP32 solver kernels.  set_tile sets both FP64
 * families, set_tile_f32 both FP32 ones.  No counterpart in the reference,
 * whose KernelParm launch shape these replace (src/QED2d.jl:10-13). */
int qed2d_ctx_set_tiles(qed2d_ctx *ctx, int q64, int q32, int cg64, int cg32);
int qed2d_ctx_get_tiles(qed2d_ctx *ctx, int tiles[4]);
/* Tile heights of the second kernel of a CG iteration (Q t and the r update)
 * of the FP64 and the FP32 solver, where they should differ from the first
 * kernel's (0: the family's height, the default; set_tiles with a nonzero
 * cg64 / cg32 resets them to 0).  get_tiles_k2 returns the heights in use.
 * No counterpart in the reference (KernelParm, src/QED2d.jl:10-13). */
int qed2d_ctx_set_tiles_k2(qed2d_ctx *ctx, int cg64, int cg32);
int qed2d_ctx_get_tiles_k2(qed2d_ctx *ctx, int tiles[2]);
/* The autotuner: times each tile height of the four families on this
 * context's lattice (scratch fields; CG families over `iters` iterations of a
 * solve that cannot converge, every pair of heights of the iteration's two
 * kernels, in two passes) and keeps the fastest; the pairs differ by a few per
 * cent, which 100 or more iterations resolve.  tiles (may be NULL)
 * receives the choice.  Tile heights change the order of the solvers' global
 * sums, so a run that must reproduce bits fixes its tiles instead. */
int qed2d_ctx_autotune(qed2d_ctx *ctx, int iters, int tiles[4]);
/* Achievable HBM bandwidth of the context's device, for the roofline beside
 * the datasheet peak: STREAM copy, triad and read over three scratch arrays of
 * nbytes each (allocated and freed here), and the copy of the same bytes as 16
 * arrays into 16 others, each block copying 256 elements of one array and
 * consecutive blocks taking consecutive arrays (the concurrent streams of a
 * 16-right-hand-side apply); gbs[4] receives GB/s counting the bytes each
 * kernel must move.  No counterpart in the reference. */
int qed2d_hbm_stream(qed2d_ctx *ctx, long nbytes, double gbs[4]);
/* dH = hfin - hini of the last trajectory of qed2d_hmc_pure / _nf2 / _rhmc on
 * this context, formed from the double-double energies before they are
 * rounded (hfin - hini of the returned doubles resolves only the ulp of H,
 * 2^-30 at 1024^2).  The Metropolis test uses exp(-(hfin - hini)) as the
 * reference (src/QED2dHMC.jl:95).  No counterpart in the reference. */
int qed2d_ctx_last_dh(const qed2d_ctx *ctx, double *dH);
/* Mixed-precision CG: a reliable update (FP64 true residual) once the FP32
 * iterated residual has dropped by delta since the last one (default 1e-3). */
int qed2d_ctx_set_reliable_delta(qed2d_ctx *ctx, double delta);
/* Where the mixed-precision CG's reliable updates are decided.  1 (default):
 * on the device -- an FP32 iteration whose residual has dropped by delta only
 * requests an update, and the update's kernels, queued at the start of every
 * chunk of iterations the host plans, run iff one is pending, so the host
 * never waits for one; the host queues each chunk before it looks at the
 * previous one.  0: the solve pauses at the request and the host queues the
 * update after reading the state (the schedule of the batched solvers, so
 * qed2d_cg_mrhs and qed2d_batch_* reproduce qed2d_cg bit for bit in this mode
 * only).  Both stop on the same FP64 true residual.  No counterpart in the
 * reference, whose CG is FP64 only (src/QED2dSolver.jl:3-105). */
int qed2d_ctx_set_reliable_mode(qed2d_ctx *ctx, int on_device);
/* Precision of the RHMC force's solves with the multi-shift force of all
 * pseudofermions (qed2d_ctx_set_multishift 1, the default).  QED2D_PREC_F64
 * (default): one FP64 multi-shift solve, as MultiCG (src/QED2dRHMC.jl:281-290).
 * QED2D_PREC_MIXED: one FP32 multi-shift solve of every pole to |r|^2 < 1e-10
 * |phi|^2 (or the solve's tolerance, if looser), then every pole's system
 * (Q^2 + mu^2) x = phi solved from that guess by one batched mixed-precision
 * CG (FP32 iterations, FP64 reliable updates) to the FP64 stopping rule; needs
 * an even L1 >= 4.  The heatbath and the acceptance step stay FP64.  No
 * counterpart in the reference, whose solvers are FP64 only. */
int qed2d_ctx_set_rhmc_precision(qed2d_ctx *ctx, int precision);
int qed2d_synchronize(qed2d_ctx *ctx);

/* ---- Dirac operator -------------------------------------------------------- */
/* gamm5Dw(so, U, si, am0, prm) kernel: so = gamma5 Dw si (src/QED2dDirac.jl:2-33) */
int qed2d_gamm5Dw(qed2d_ctx *ctx, double *so, const double *U, const double *si, double am0);
/* FP32 form of the same operator (float2 fields, same layout). */
int qed2d_gamm5Dw_f32(qed2d_ctx *ctx, float *so, const float *U, const float *si, float am0);
/* nrhs right-hand sides sharing U; RHS j at si + j*4V doubles. */
int qed2d_gamm5Dw_mrhs(qed2d_ctx *ctx, double *so, const double *U, const double *si, int nrhs,
                       double am0);
/* CG (v1 semantics, operator Q^2 [+ 0]) on nrhs right-hand sides sharing U, all
 * in one launch per kernel; x and b are nrhs fields back to back ((k, 2, L2, L1)),
 * iters (may be NULL) receives the iterations per right-hand side.  Each
 * solution equals qed2d_cg on that right-hand side with the same tiles, bit
 * for bit (mixed precision: with qed2d_ctx_set_reliable_mode(ctx, 0)).  No
 * counterpart in the reference, which solves one right-hand side per call
 * (src/QED2dSolver.jl:3-45). */
int qed2d_cg_mrhs(qed2d_ctx *ctx, double *x, const double *U, const double *b, int nrhs,
                  double am0, int maxiter, double eps, int precision, int *iters);
int qed2d_gamm5Dw_mrhs_f32(qed2d_ctx *ctx, float *so, const float *U, const float *si, int nrhs,
                           float am0);
/* gamm5(so, si, am0, prm) kernel (src/QED2dDirac.jl:130-146) */
int qed2d_gamm5(qed2d_ctx *ctx, double *so, const double *si);
/* Dw(so, U, si, am0, prm, kprm) (src/QED2dDirac.jl:157-169) */
int qed2d_Dw(qed2d_ctx *ctx, double *so, const double *U, const double *si, double am0);
/* gamm5Dw_sqr(so, U, si, am0, prm, kprm) (src/QED2dDirac.jl:35-46) */
int qed2d_gamm5Dw_sqr(qed2d_ctx *ctx, double *so, const double *U, const double *si, double am0);
/* gamm5Dw_sqr_msq(so, tmp, U, si, am0, prm, kprm) (src/QED2dDirac.jl:69-81); tmp may be NULL */
int qed2d_gamm5Dw_sqr_msq(qed2d_ctx *ctx, double *so, double *tmp, const double *U,
                          const double *si, double am0);
/* gamm5Dw_sqr_musq(so, tmp, U, si, am0, mu, prm, kprm) (src/QED2dDirac.jl:83-95) */
int qed2d_gamm5Dw_sqr_musq(qed2d_ctx *ctx, double *so, double *tmp, const double *U,
                           const double *si, double am0, double mu);
/* gamm5Dw_sqr_sqr(so, U, si, am0, prm, kprm) (src/QED2dDirac.jl:48-67) */
int qed2d_gamm5Dw_sqr_sqr(qed2d_ctx *ctx, double *so, const double *U, const double *si,
                          double am0);
/* gamm5Dw_sqr_sqr_musq(so, tmp, U, si, am0, mu, prm, kprm) (src/QED2dDirac.jl:97-117) */
int qed2d_gamm5Dw_sqr_sqr_musq(qed2d_ctx *ctx, double *so, double *tmp, const double *U,
                               const double *si, double am0, double mu);

/* ---- forces, link update, observables -------------------------------------- */
/* tr_dQwdU(frc, U, X, g5DwX, prm) kernel (src/QED2dDirac.jl:172-199) */
int qed2d_tr_dQwdU(qed2d_ctx *ctx, double *frc, const double *U, const double *X,
                   const double *g5DwX);
/* krnl_force!(frc1, frc2, U, prm) kernel (src/QED2dAction.jl:34-53) */
int qed2d_krnl_force(qed2d_ctx *ctx, double *frc1, double *frc2, const double *U);
/* updt_fld!(U, mom, eps) kernel (src/QED2dAction.jl:55-64) */
int qed2d_updt_fld(qed2d_ctx *ctx, double *U, const double *mom, double eps);
/* mom .+= eps (frc1 + frc2 + frc) fused with the force kernels
 * (src/QED2dHMC.jl:463-478): gauge force, plus the Nf=2 fermion force of
 * (X, g5DwX) when X is not NULL. */
int qed2d_force_mom_update(qed2d_ctx *ctx, double *mom, const double *U, const double *X,
                           const double *g5DwX, double eps);
/* plaquette!(plx, U, prm) kernel (src/QED2dAction.jl:2-16) */
int qed2d_plaquette_field(qed2d_ctx *ctx, double *plx, const double *U);
/* qtop!(plx, U, prm) kernel (src/QED2dAction.jl:18-32) */
int qed2d_qtop_field(qed2d_ctx *ctx, double *plx, const double *U);
/* Action(U, prm, kprm) (src/QED2dHMC.jl:206-215) */
int qed2d_action(qed2d_ctx *ctx, const double *U, double *S);
/* Plaquette(U, prm, kprm) (src/QED2dHMC.jl:217-226) */
int qed2d_plaquette(qed2d_ctx *ctx, const double *U, double *P);
/* Qtop(U, prm, kprm) (src/QED2dHMC.jl:228-237) */
int qed2d_qtop(qed2d_ctx *ctx, const double *U, double *Q);
/* Hamiltonian(mom, U, prm, kprm) (src/QED2dHMC.jl:198-203) */
int qed2d_hamiltonian(qed2d_ctx *ctx, const double *mom, const double *U, double *H);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): Hamiltonian,
 * Action, Plaquette and Qtop of one L1 x L2 field (as qed2d_hamiltonian,
 * qed2d_action, qed2d_plaquette and qed2d_qtop, same bits) with no context and
 * no allocation.  mom may be NULL (then H = S_g).  out is a DEVICE buffer of 4
 * doubles: H, S_g, <P>, Q.  Asynchronous on `stream` (a hipStream_t, NULL =
 * default stream).  The reduction's scratch is a fixed 48 KiB array per device
 * (enough for any lattice); every call on a device shares it, so calls that
 * overlap on one GPU race.  The first call on a device resolves the array's
 * address; make it before a graph capture starts.  Lattices beyond 2^31 sites
 * or below 2 x 2 return QED2D_EINVAL. */
int qed2d_hamiltonian_noalloc(int L1, int L2, double beta, const double *mom, const double *U,
                              double *out, void *stream);
/* CUDA.dot(a, b) and CUDA.mapreduce(abs2, +, a) on spinors (src/QED2dSolver.jl:13,21) */
int qed2d_dot(qed2d_ctx *ctx, const double *a, const double *b, double *re, double *im);
int qed2d_norm2(qed2d_ctx *ctx, const double *a, double *n2);

/* ---- solvers ------------------------------------------------------------------
 * CG(so, U, si, am0, maxiter, eps, A, prm, kprm)            (src/QED2dSolver.jl:3-45)
 * CG(so, U, si, am0, mu, maxiter, eps, A, prm, kprm;
 *    so_as_guess)                                          (src/QED2dSolver.jl:49-105)
 * op selects A; stop when |r|^2 < eps |b|^2 (eps is a SQUARED relative
 * residual, as in the reference).  precision QED2D_PREC_MIXED runs FP32
 * iterations with FP64 reliable updates and stops on the same FP64 true
 * residual.  *iters receives the iteration count (0 when the guess already
 * satisfies the tolerance, as the reference).  Returns QED2D_ENOCONV after
 * maxiter iterations without convergence. */
int qed2d_cg(qed2d_ctx *ctx, double *so, const double *U, const double *si, double am0,
             double mu, int op, int maxiter, double eps, int so_as_guess, int precision,
             int *iters);
/* MultiCG(so, U, si, am0, maxiter, eps, A, rprm, prm, kprm) (src/QED2dRHMC.jl:281-290):
 * so = (1 + sum_j rho_j (Q^2 + mu_j^2)^-1) si. */
int qed2d_multicg(qed2d_ctx *ctx, double *so, const double *U, const double *si, double am0,
                  int n, const double *mu, const double *rho, int maxiter, double eps);
/* Multi-shift CG: (Q^2 + mu_i^2) x_i = b for i < n <= 32 on one Krylov
 * sequence (Jegerlehner, hep-lat/9612014).  x is a host array of n device
 * pointers.  Replaces the n sequential CG solves of MultiCG / R
 * (src/QED2dRHMC.jl:264-267,284-287) and of the RHMC force
 * (src/QED2dHMC.jl:528-546); each shift stops on the reference's rule
 * |r_i|^2 < eps |b|^2. */
int qed2d_multishift_cg(qed2d_ctx *ctx, double *const *x, const double *U, const double *b,
                        double am0, int n, const double *mu, int maxiter, double eps, int *iters);
/* How MultiCG, R and the RHMC force solve their shifted systems:
 * 1 (default): multi-shift CG, and in an RHMC trajectory all pseudofermions'
 *    solves together (one launch per kernel over the right-hand sides): the
 *    force's and the acceptance's multi-shift solves, and the heatbath's
 *    shifted CG solves pole by pole;
 * 2: multi-shift CG, one pseudofermion after another;
 * 0: the reference's sequential (warm-started) CG solves.
 * Modes 1 and 2 give the same bits. */
int qed2d_ctx_set_multishift(qed2d_ctx *ctx, int mode);
/* R(so, U, si, am0, maxiter, eps, A, rprm, prm, kprm) (src/QED2dRHMC.jl:261-271) */
int qed2d_R(qed2d_ctx *ctx, double *so, const double *U, const double *si, double am0, int n,
            const double *mu, const double *rho, double A, double r_b, int maxiter, double eps);
/* generate_pseudofermion!(F, U, X, am0, CGmaxiter, CGtol, prm, kprm, rprm)
 * (src/QED2dRHMC.jl:298-319) */
int qed2d_generate_pseudofermion(qed2d_ctx *ctx, double *F, const double *U, const double *X,
                                 double am0, int n, const double *mu, const double *nu,
                                 int maxiter, double tol);
/* power_method(U, am0, prm, kprm; iter) (src/QED2dRHMC.jl:135-170); b0 is the
 * starting vector (the reference draws it at :137). */
int qed2d_power_method(qed2d_ctx *ctx, const double *U, const double *b0, double am0, int iter,
                       double *lambda_min, double *lambda_max);

/* LuscherZp(Z, p, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm)
 * (src/QED2dRHMC.jl:189-201; p = 1 is LuscherZ, :174-186): Z = (Q^2 R^2 - 1)^p X.
 * Z must not alias X.  p = 0 leaves Z unchanged. */
int qed2d_luscher_z(qed2d_ctx *ctx, double *Z, int p, const double *U, const double *X,
                    double am0, int n, const double *mu, const double *rho, double A, double r_b,
                    int maxiter, double tol);
/* reweighting_factor(U, am0, CGmaxiter, CGtol, prm, kprm, rprm)
 * (src/QED2dRHMC.jl:210-239): W_N from N Gaussian fields and a Taylor expansion of
 * order `taylor` (RHMCParm.reweighting_N / reweighting_Taylor).  The j-th field is
 * drawn from the (seed, chain) stream at trajectory `traj` (it replaces CUDA.randn
 * at :218).  The product over several fermions (:242-251) is left to the caller. */
int qed2d_reweighting_factor(qed2d_ctx *ctx, const double *U, double am0, int n,
                             const double *mu, const double *rho, double A, double r_b, int N,
                             int taylor, int maxiter, double tol, uint64_t seed, uint32_t chain,
                             uint32_t traj, double *W);

/* ---- batched ensembles ------------------------------------------------------
 * Many independent chains (lattices of different sizes and couplings) advanced
 * together; no counterpart in the reference, which runs one lattice per
 * process (main/hmc-nf2.jl).  The links of all chains live in one device arena
 * owned by the batch: chain c's links are a reference-layout [L1,L2,2] field at
 * qed2d_batch_links(b, c) (site offset qed2d_batch_offset).  Every CG
 * iteration, Q apply, force + momentum update and link update of a call is
 * one launch over all chains; each chain's trajectory is HMC! for Nf=2
 * (src/QED2dHMC.jl:64-113) drawn from the streams of (seed[c], chain c, traj),
 * the same as qed2d_hmc_nf2(..., seed[c], c, traj, ...) on that lattice. */
typedef struct qed2d_batch qed2d_batch;
int qed2d_batch_create(qed2d_batch **b, int nchain, const int *L1, const int *L2,
                       const double *beta, int bc_flags, void *stream);
void qed2d_batch_destroy(qed2d_batch *b);
const char *qed2d_batch_last_error(const qed2d_batch *b);
int qed2d_batch_size(const qed2d_batch *b, int *nchain, long *sites);
int qed2d_batch_offset(const qed2d_batch *b, int chain, long *site_offset);
int qed2d_batch_links(qed2d_batch *b, int chain, double **U);
/* Copies of chain c's links out of / into the arena (host or device buffer,
 * reference layout; synchronous). */
int qed2d_batch_get_links(qed2d_batch *b, int chain, double *dst);
int qed2d_batch_set_links(qed2d_batch *b, int chain, const double *src);
/* Tile heights of the batch's FP64 apply, FP64 and FP32 solver kernels
 * (0 leaves one unchanged; a nonzero cg64 / cg32 also resets the second CG
 * kernel's height to follow it).  The defaults are the tallest heights (16 / 8
 * rows) that leave a solver launch of at least 1024 blocks over all chains.
 * set_tiles_k2: the second CG kernel's heights (0: the family's).  get_tiles:
 * q64, cg64, cg32, cg64 of K2, cg32 of K2 in use.  autotune: times every apply
 * height and every (K1, K2) pair of the FP64 and FP32 CG (the pairs in two
 * passes) on the batch's own lattices and links over `iters` iterations (the
 * pseudofermion fields serve as scratch) and keeps the fastest.  No
 * counterpart in the reference
 * (KernelParm, src/QED2d.jl:10-13). */
int qed2d_batch_set_tiles(qed2d_batch *b, int q64, int cg64, int cg32);
int qed2d_batch_set_tiles_k2(qed2d_batch *b, int cg64, int cg32);
int qed2d_batch_get_tiles(qed2d_batch *b, int tiles[5]);
int qed2d_batch_autotune(qed2d_batch *b, int iters, int tiles[5]);
/* Mixed-precision CG of a batch: on_device != 0 (the default) lets a chain
 * whose FP32 residual has dropped by the reliable-update factor request the
 * update and iterate on until the device performs it at the next chunk start;
 * 0 pauses the chain until then, which reproduces qed2d_cg with
 * qed2d_ctx_set_reliable_mode(ctx, 0) on that chain bit for bit.  No
 * counterpart in the reference. */
int qed2d_batch_set_reliable_mode(qed2d_batch *b, int on_device);
/* on != 0: the batch's CG solves queue each chunk of iterations (its reliable-
 * update slot and K1/K2 launches) as one HIP graph launch, captured on first
 * use per chunk shape and kept until qed2d_batch_destroy; the same kernels and
 * bits as the direct launches (default off).  No counterpart in the reference. */
int qed2d_batch_set_graphs(qed2d_batch *b, int on);
/* Hot start of every chain: link angles uniform in [-pi, pi) from (seed[c], c). */
int qed2d_batch_hot_start(qed2d_batch *b, const uint64_t *seed);
/* (Q^2) x = src on every chain from x = 0 (CG v1 semantics per chain); x and
 * src are arena-shaped (2 x total sites complex).  iters[c] per chain. */
int qed2d_batch_cg(qed2d_batch *b, double *x, const double *src, const double *am0,
                   int maxiter, double tol, int precision, int *iters);
/* A set of nrhs right-hand sides of one L1 x L2 lattice that share the caller's
 * links (the chain-set machinery with every chain's links at the same field):
 * qed2d_batch_cg_shared solves (Q^2) x_j = src_j for all of them, one launch
 * per kernel; x and src hold nrhs reference-layout fields back to back. */
int qed2d_batch_create_rhs(qed2d_batch **b, int L1, int L2, int nrhs, double beta,
                           int bc_flags, void *stream);
int qed2d_batch_cg_shared(qed2d_batch *b, const double *U, double *x, const double *src,
                          const double *am0, int maxiter, double tol, int precision, int *iters);
/* One Nf=2 trajectory of every chain; am0, seed, hini, hfin, accepted, cg_iters
 * hold one entry per chain (cg_iters may be NULL). */
int qed2d_batch_hmc_nf2(qed2d_batch *b, const double *am0, double eps, int ns, int integrator,
                        int maxiter, double tol, int precision, double acc_tol,
                        const uint64_t *seed, uint32_t traj, double *hini, double *hfin,
                        int *accepted, int *cg_iters);
/* Per-chain dH of the last trajectory from the double-double energies (see
 * qed2d_ctx_last_dh). */
int qed2d_batch_last_dh(const qed2d_batch *b, double *dH);
/* Per-chain plaquette, topological charge and gauge action (any may be NULL). */
int qed2d_batch_observables(qed2d_batch *b, double *plaquette, double *qtop, double *action);

/* ---- Zolotarev coefficients (host only) ---------------------------------------
 * get_rhmc_params(n, r_a, r_b) (src/QED2dRHMC.jl:93-110) -> RHMCParm fields
 * (src/QED2d.jl:15-26).  rho, mu, nu hold n doubles each. */
int qed2d_zolotarev(int n, double r_a, double r_b, double *eps, double *A, double *delta,
                    double *rho, double *mu, double *nu);

/* ---- random fields -------------------------------------------------------------
 * Counter-based Philox4x32-10 streams keyed by (seed, chain) and indexed by
 * trajectory; they replace CUDA.randn / rand() at src/QED2dHMC.jl:25,42,69-70,
 * 95,124,138 (CURAND streams are not reproducible outside CUDA). */
int qed2d_fill_momenta(qed2d_ctx *ctx, double *mom, uint64_t seed, uint32_t chain,
                       uint32_t traj);
int qed2d_fill_pseudofermion(qed2d_ctx *ctx, double *X, uint64_t seed, uint32_t chain,
                             uint32_t traj, uint32_t flavour);
int qed2d_hot_start(qed2d_ctx *ctx, double *U, uint64_t seed, uint32_t chain);
double qed2d_metropolis_uniform(uint64_t seed, uint32_t chain, uint32_t traj);

/* ---- molecular dynamics ---------------------------------------------------------
 * OMF4!(mom, U, eps, ns, prm, kprm)                 (src/QED2dHMC.jl:239-305) */
int qed2d_omf4_pure(qed2d_ctx *ctx, double *mom, double *U, double eps, int ns);
/* leapfrog!(mom, U, X, F, g5DX, am0, eps, ns, ...)  (src/QED2dHMC.jl:426-449)
 * OMF4!(mom, U, X, F, g5DX, am0, eps, ns, ...)      (src/QED2dHMC.jl:307-422)
 * X receives the last solve; g5DX the last Q X.  *cg_iters: total. */
int qed2d_md_nf2(qed2d_ctx *ctx, int integrator, double *mom, double *U, double *X,
                 const double *F, double *g5DX, double am0, double eps, int ns, int maxiter,
                 double tol, int precision, int *cg_iters);

/* ---- HMC trajectories --------------------------------------------------------------
 * HMC!(U, eps, ns, acc, prm, kprm; integrator, qzero)         (src/QED2dHMC.jl:19-60) */
int qed2d_hmc_pure(qed2d_ctx *ctx, double *U, double eps, int ns, int integrator, int qzero,
                   uint64_t seed, uint32_t chain, uint32_t traj, double *hini, double *hfin,
                   int *accepted);
/* HMC!(U, am0, eps, ns, acc, CGmaxiter, tol, prm, kprm; integrator, qzero)
 * (src/QED2dHMC.jl:64-113).  precision selects the MD-force solver.
 * acc_tol <= 0: the acceptance step reuses the last MD solve, as the
 * reference (:82).  acc_tol > 0: the acceptance step re-solves in FP64 to
 * that squared relative residual, starting from the last MD solve, and
 * evaluates S_f = 2 Re(X^dag F) - |Q X|^2 (error second order in the
 * residual). */
int qed2d_hmc_nf2(qed2d_ctx *ctx, double *U, double am0, double eps, int ns, int integrator,
                  int maxiter, double tol, int qzero, int precision, double acc_tol,
                  uint64_t seed, uint32_t chain, uint32_t traj, double *hini, double *hfin,
                  int *accepted, int *cg_iters);
/* HMC!(U, am0::Array, eps, ns, acc, CGmaxiter, tol, prm, kprm, rprm::Array;
 *      integrator, qzero, reversibility_test)                (src/QED2dHMC.jl:116-195)
 * nf flavours; flavour j has npole[j] poles mu[j][], nu[j][], rho[j][].
 * integrator QED2D_OMF4 is an extension: the reference's MD! has no OMF4
 * method for RHMC (src/QED2dHMC.jl:12-16). */
int qed2d_hmc_rhmc(qed2d_ctx *ctx, double *U, const double *am0, int nf, const int *npole,
                   const double *const *mu, const double *const *nu, const double *const *rho,
                   double eps, int ns, int integrator, int maxiter, double tol, int qzero,
                   int reversibility_test, uint64_t seed, uint32_t chain, uint32_t traj,
                   double *hini, double *hfin, int *accepted);

/* ---- gauge-configuration files -----------------------------------------------------
 * save_gauge(U, file, prm) / load_gauge(U, file, prm) (src/QED2dLoadSave.jl:1-19):
 * text, one link per line "re<TAB>im" in column-major order.  U is a device
 * field; the host-only forms take a host array of nlinks = 2 L1 L2 complex numbers. */
int qed2d_save_gauge(qed2d_ctx *ctx, const double *U, const char *path);
int qed2d_load_gauge(qed2d_ctx *ctx, double *U, const char *path);
int qed2d_write_links_text(const double *U_host, long nlinks, const char *path);
int qed2d_read_links_text(double *U_host, long nlinks, const char *path);

/* ---- measurement -----------------------------------------------------------------------
 * HIP-event timing of the library's own kernels, per kernel class, on the
 * context's stream (no counterpart in the reference, which only has @time). */
#define QED2D_PROF_CG_K1_F64 0    /* CG kernel 1, FP64: p = r + beta p, t = Q p, |t|^2 */
#define QED2D_PROF_CG_K2_F64 1    /* CG kernel 2, FP64: Ap = Q t, x/r update, |r|^2 */
#define QED2D_PROF_CG_K1_F32 2
#define QED2D_PROF_CG_K2_F32 3
#define QED2D_PROF_Q_F64 4        /* plain gamm5Dw apply */
#define QED2D_PROF_FORCE_MOM 5    /* fused force + momentum update */
#define QED2D_PROF_UPDT_FLD 6
#define QED2D_PROF_RELIABLE 7     /* reliable-update kernels of the mixed CG */
#define QED2D_PROF_OTHER 8
#define QED2D_PROF_NCLASS 9
/* every = 0: off; every = N: time every N-th launch of each class. */
int qed2d_profile_enable(qed2d_ctx *ctx, int every);
int qed2d_profile_reset(qed2d_ctx *ctx);
int qed2d_profile_read(qed2d_ctx *ctx, int cls, long *count, double *total_ms);

/* ---- fields ----------------------------------------------------------------------------
 * unfold_fld! / unfold_reflect! (src/QED2dFields.jl:1-50): Uout is (2L1)x(2L2). */
int qed2d_unfold(qed2d_ctx *ctx, double *Uout, const double *Uin, int reflect);

#ifdef __cplusplus
}
#endif
#endif
