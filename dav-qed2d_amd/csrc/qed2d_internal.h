// qed2d_internal.h -- host-side declarations shared between the kernel
// translation unit (qed2d_kernels.hip) and the orchestration / C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qed2d_device.h"

namespace qed2d {

// Device-resident solver state.  Kernels read the scalars they need and the
// last block of each reduction writes the new ones, so a whole CG iteration
// runs without a host round trip.
struct SolverState {
    double norm;    // |r|^2 entering the current iteration
    double tol;     // eps * |b|^2
    double beta;    // coefficient of the next search-direction update
    double tt;      // |Q p|^2
    double pp;      // |p|^2
    double err;     // |r|^2 after the last update
    double mu2;     // shift mu^2 of the operator Q^2 + mu^2
    double rmax2;   // largest |r|^2 since the last reliable update
    double delta2;  // reliable-update trigger (squared residual reduction)
    double bnorm;   // |b|^2
    double eps;     // requested squared relative residual
    int iter;       // iterations completed
    int maxiter;
    int status;     // 0 running, 1 converged, 2 iteration limit reached
    int need_ru;    // mixed precision: a reliable update is due
    int n_ru;       // reliable updates performed
    int pad;
    unsigned ticket[8];
};

// Multi-shift CG state (all shifts sigma_i = mu_i^2 of (Q^2 + sigma_i) x_i = b on
// one Krylov sequence).  Shift i is active 1 while iterating, 2 once converged
// with its last x update still pending, 0 when finished.
constexpr int kMaxShift = 32;
struct MultiShiftState {
    double rr;          // |r_k|^2
    double tt, pp;      // |Q p_k|^2, |p_k|^2
    double alpha_prev;  // alpha_{k-1}
    double beta;        // beta_k (p_k = r_k + beta_k p_{k-1})
    double sigma_b;     // base shift (the smallest)
    double tol, bnorm, eps;
    int iter, maxiter, status, nshift;
    double s[kMaxShift];         // sigma_i - sigma_b
    double zeta[kMaxShift];      // zeta_k
    double zeta_prev[kMaxShift]; // zeta_{k-1}
    double alpha_s[kMaxShift];   // alpha^sigma_{k-1} (pending x update)
    double beta_s[kMaxShift];    // beta^sigma_k
    int active[kMaxShift];
    int conv_iter[kMaxShift];
    double2 *x[kMaxShift];       // solution of shift i
    double2 *ps[kMaxShift];      // search direction of shift i
    unsigned ticket[4];
};

// Scalars of a reduction that the host reads back.
struct RedOut {
    double v[8];
    unsigned ticket[8];
};

enum { kBlock = 256 };

// Launch-shape description of the marching stencil kernels.
struct MarchGrid {
    int bx, by;          // threads of a block along i1 and along i2 (bx * by = kBlock)
    int nbx, nby, rows;  // blocks along i1, blocks along i2, rows marched per thread
    int nblocks() const { return nbx * nby; }
};
MarchGrid march_grid(const Latt &L, int rows);

// ---- launches (all asynchronous on `st`) ---------------------------------
void launch_q_apply_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                        double am0, int rows, hipStream_t st);
void launch_q_apply_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so,
                        float am0, int rows, hipStream_t st);
void launch_q_apply_mrhs_f64(const Latt &L, const double2 *U, const double2 *si, double2 *so,
                             int nrhs, double am0, hipStream_t st);
void launch_q_apply_mrhs_f32(const Latt &L, const float2 *U, const float2 *si, float2 *so,
                             int nrhs, float am0, hipStream_t st);
void launch_gamm5(const Latt &L, const double2 *si, double2 *so, hipStream_t st);
// so = Q si + mu^2 x  (x may alias nothing; used for Q^2 + mu^2 with si = Q x)
void launch_q_shift_f64(const Latt &L, const double2 *U, const double2 *si, const double2 *x,
                        double2 *so, double am0, double mu2, int rows, hipStream_t st);

// CG (T = double or float).  part: per-block partial sums (>= 2*nblocks).
void launch_cg_init_f64(const Latt &L, const double2 *b, double2 *x, double2 *r, double2 *p0,
                        SolverState *s, double *part, hipStream_t st);
// r = b - (Q t + mu^2 x) with t = Q x already computed; p0 = 0; sets norm, bnorm, tol.
void launch_cg_guess_f64(const Latt &L, const double2 *U, const double2 *b, const double2 *x,
                         const double2 *t, double2 *r, double2 *p0, double am0, SolverState *s,
                         double *part, int rows, hipStream_t st);
// p0 / p1: the two search-direction buffers, used alternately by iteration
// parity on the device.
void launch_cg_k1_f64(const Latt &L, const double2 *U, const double2 *r, double2 *p0,
                      double2 *p1, double2 *t, double am0, SolverState *s, double *part,
                      int rows, hipStream_t st);
void launch_cg_k2_f64(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                      const double2 *p1, double2 *x, double2 *r, double am0, SolverState *s,
                      double *part, int rows, int mixed, hipStream_t st);
void launch_cg_k1_f32(const Latt &L, const float2 *U, const float2 *r, float2 *p0, float2 *p1,
                      float2 *t, float am0, SolverState *s, double *part, int rows,
                      hipStream_t st);
void launch_cg_k2_f32(const Latt &L, const float2 *U, const float2 *t, const float2 *p0,
                      const float2 *p1, float2 *x, float2 *r, float am0, SolverState *s,
                      double *part, int rows, int mixed, hipStream_t st);
// Mixed precision helpers.
void launch_cg_init_mixed(const Latt &L, const double2 *b, float2 *rs, float2 *xs, float2 *p0,
                          SolverState *s, double *part, hipStream_t st);
void launch_ru_accum(const Latt &L, double2 *x, float2 *xs, const SolverState *s,
                     hipStream_t st);
// rs = float(b - (Q t + mu^2 x)), t = Q x.  initial != 0: first residual of a
// solve (beta = 0, sets bnorm and tol); otherwise a reliable update.
void launch_ru_residual(const Latt &L, const double2 *U, const double2 *b, const double2 *t,
                        const double2 *x, float2 *rs, float2 *p0, double am0, SolverState *s,
                        double *part, int rows, int initial, hipStream_t st);
void launch_to_f32(const double2 *in, float2 *out, long n, hipStream_t st);

// Multi-shift CG: init (r = b, x_i = 0, p_i = 0, p0 = 0, |b|^2), the two
// per-iteration kernels and the pending final x updates.
void launch_ms_init(const Latt &L, const double2 *b, double2 *r, double2 *p0,
                    MultiShiftState *s, double *part, int nshift, hipStream_t st);
void launch_ms_k1(const Latt &L, const double2 *U, const double2 *r, double2 *p0, double2 *p1,
                  double2 *t, double am0, MultiShiftState *s, double *part, int nshift, int rows,
                  hipStream_t st);
void launch_ms_k2(const Latt &L, const double2 *U, const double2 *t, const double2 *p0,
                  const double2 *p1, double2 *r, double am0, MultiShiftState *s, double *part,
                  int nshift, int rows, hipStream_t st);
void launch_ms_final(const Latt &L, MultiShiftState *s, int nshift, hipStream_t st);

// Reductions (deterministic; results in out->v[]).
void launch_norm2(const double2 *a, long n, RedOut *out, double *part, hipStream_t st);
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
void launch_negate(double *a, long n, hipStream_t st);

// RNG fills.
void launch_fill_momenta(const Latt &L, double *mom, uint32_t k0, uint32_t k1, uint32_t traj,
                         hipStream_t st);
void launch_fill_pseudofermion(const Latt &L, double2 *X, uint32_t k0, uint32_t k1,
                               uint32_t traj, uint32_t flavour, hipStream_t st);
void launch_hot_start(const Latt &L, double2 *U, uint32_t k0, uint32_t k1, hipStream_t st);
void launch_unfold(const Latt &L, const double2 *Uin, double2 *Uout, int reflect,
                   hipStream_t st);

}  // namespace qed2d
