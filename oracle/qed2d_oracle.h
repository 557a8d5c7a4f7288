/*
 * qed2d_oracle.h -- FP64 CPU restatement of the QED2d molecular-dynamics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the FP64 restatement that the HIP
 * library (libqed2d_hip) is compared against.  It is loaded by tests/, by the smoke
 * check in __graft_entry__.py and by the cpu_baseline leg of bench.py, and by
 * nothing else.  No product code path may call it.
 *
 * Every function restates the reference Julia package dalbandea/Dav-QED2d
 * (module QED2d) kernel by kernel, in the same operation order, unfused, and
 * cites the file:line it follows.  Fields use the reference's own layout:
 * Julia column-major [L1,L2,2], i.e. C index ((c*L2 + i2)*L1 + i1), complex
 * numbers interleaved (re, im).
 *
 * Pinning (what checks each function; all in tests/test_oracle_*.py):
 *   gamm5Dw/gamm5/Dw/sqr*   dense gamma-matrix Dirac matrix (numpy), free-field
 *                           plane-wave eigenvalues, gamma5-hermiticity, gauge
 *                           covariance
 *   tr_dQwdU, krnl_force    central finite differences of the actions
 *   updt_fld, plaquette,    closed forms (cold start, constant-flux fields)
 *   qtop
 *   cg_v1/v2, multicg, R    dense solves (numpy.linalg.solve)
 *   zolotarev               mpmath golden vectors (tests/golden) and the
 *                           equioscillation bound
 *   power_method            dense eigenvalues of Q^2
 *   generate_pseudofermion  X^dag X = phi^dag (1+sum rho (Q^2+mu^2)^-1) phi
 *   integrators / HMC       reversibility, dH ~ eps^2 (leapfrog), eps^4 (OMF4),
 *                           exact energy conservation of the pure-gauge flow
 *   philox4x32_10           Random123 known-answer vectors
 *   gauss/uniform           moments and Kolmogorov-Smirnov tests
 */
#ifndef QED2D_ORACLE_H
#define QED2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fermion boundary flags: bit 0 set = antiperiodic in direction 2. */
#define ORC_BC_APERIODIC_2 1

/* ---- random numbers (counter-based Philox streams) ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_det_log(double u);
void orc_det_sincos2pi(double u, double *s, double *c);
void orc_gauss_pair(const uint32_t ctr[4], const uint32_t key[2], double *g1, double *g2);
double orc_uniform(const uint32_t ctr[4], const uint32_t key[2]);
void orc_fill_momenta(double *mom, int L1, int L2, uint64_t seed, uint32_t chain, uint32_t traj);
void orc_fill_cgauss(double *X, int L1, int L2, uint64_t seed, uint32_t chain, uint32_t traj,
                     uint32_t purpose, uint32_t sub0);
void orc_fill_pseudofermion(double *X, int L1, int L2, uint64_t seed, uint32_t chain,
                            uint32_t traj, uint32_t flavour);
double orc_metropolis_uniform(uint64_t seed, uint32_t chain, uint32_t traj);
void orc_hot_start(double *U, int L1, int L2, uint64_t seed, uint32_t chain);

/* ---- Dirac operator ------------------------------------------------------ */
void orc_gamm5Dw(double *so, const double *U, const double *si, double am0,
                 int L1, int L2, int bc);
void orc_gamm5(double *so, const double *si, int L1, int L2);
void orc_Dw(double *so, const double *U, const double *si, double am0, int L1, int L2, int bc);
void orc_gamm5Dw_sqr(double *so, const double *U, const double *si, double am0,
                     int L1, int L2, int bc);
void orc_gamm5Dw_sqr_msq(double *so, double *tmp, const double *U, const double *si,
                         double am0, int L1, int L2, int bc);
void orc_gamm5Dw_sqr_musq(double *so, double *tmp, const double *U, const double *si,
                          double am0, double mu, int L1, int L2, int bc);
void orc_gamm5Dw_sqr_sqr(double *so, const double *U, const double *si, double am0,
                         int L1, int L2, int bc);
void orc_gamm5Dw_sqr_sqr_musq(double *so, double *tmp, const double *U, const double *si,
                              double am0, double mu, int L1, int L2, int bc);

/* ---- forces, updates, observables --------------------------------------- */
void orc_tr_dQwdU(double *frc, const double *U, const double *X, const double *Y,
                  int L1, int L2, int bc);
void orc_krnl_force(double *frc1, double *frc2, const double *U, double beta, int L1, int L2);
void orc_updt_fld(double *U, const double *mom, double eps, int L1, int L2);
void orc_plaquette_field(double *plx, const double *U, int L1, int L2);
void orc_qtop_field(double *plx, const double *U, int L1, int L2);
double orc_action(const double *U, double beta, int L1, int L2);
double orc_plaquette(const double *U, int L1, int L2);
double orc_qtop(const double *U, int L1, int L2);
double orc_hamiltonian(const double *mom, const double *U, double beta, int L1, int L2);
void orc_dot(const double *a, const double *b, long n, double *re, double *im);
double orc_norm2(const double *a, long n);

/* ---- solvers -------------------------------------------------------------
 * op: 0 = gamm5Dw_sqr_msq (Q^2), 1 = gamm5Dw_sqr_musq (Q^2 + mu^2).
 * Return: iteration count; *status = 0 converged, 1 not converged.        */
int orc_cg(double *so, const double *U, const double *si, double am0, double mu, int op,
           int maxiter, double eps, int so_as_guess, int L1, int L2, int bc, int *status);
int orc_multicg(double *so, const double *U, const double *si, double am0, int n,
                const double *mu, const double *rho, int maxiter, double eps,
                int L1, int L2, int bc);
int orc_R(double *so, const double *U, const double *si, double am0, int n,
          const double *mu, const double *rho, double A, double r_b, int maxiter,
          double eps, int L1, int L2, int bc);
int orc_generate_pseudofermion(double *F, const double *U, const double *X, double am0,
                               int n, const double *mu, const double *nu, int maxiter,
                               double tol, int L1, int L2, int bc);
void orc_power_method(const double *U, const double *b0, double am0, int iter, int L1,
                      int L2, int bc, double *lmin, double *lmax);

/* ---- Zolotarev (own elliptic functions) --------------------------------- */
double orc_ellipk(double m);
void orc_ellip_sncn(double u, double m, double *sn, double *cn);
/* out: rho[n], mu[n], nu[n]; scalars via pointers. */
void orc_zolotarev(int n, double r_a, double r_b, double *eps, double *A, double *delta,
                   double *rho, double *mu, double *nu);

/* ---- molecular dynamics and HMC ----------------------------------------- */
typedef struct {
    int L1, L2, bc;
    double beta;
} orc_latt;

int orc_update_momenta_nf2(double *mom, const double *U, double *X, const double *F,
                           double *g5DX, double am0, double eps, int maxiter, double tol,
                           const orc_latt *p);
int orc_leapfrog_nf2(double *mom, double *U, double *X, const double *F, double *g5DX,
                     double am0, double eps, int ns, int maxiter, double tol,
                     const orc_latt *p);
int orc_omf4_nf2(double *mom, double *U, double *X, const double *F, double *g5DX,
                 double am0, double eps, int ns, int maxiter, double tol, const orc_latt *p);
void orc_omf4_pure(double *mom, double *U, double eps, int ns, const orc_latt *p);
void orc_leapfrog_pure(double *mom, double *U, double eps, int ns, const orc_latt *p);

int orc_update_momenta_rhmc(double *mom, const double *U, double *X, double *const *F,
                            double *g5DX, const double *am0, int nf, const int *npole,
                            double *const *mu, double *const *rho, double eps, int maxiter,
                            double tol, const orc_latt *p);
int orc_leapfrog_rhmc(double *mom, double *U, double *X, double *const *F, double *g5DX,
                      const double *am0, int nf, const int *npole, double *const *mu,
                      double *const *rho, double eps, int ns, int maxiter, double tol,
                      const orc_latt *p);

int orc_omf4_rhmc(double *mom, double *U, double *X, double *const *F, double *g5DX,
                  const double *am0, int nf, const int *npole, double *const *mu,
                  double *const *rho, double eps, int ns, int maxiter, double tol,
                  const orc_latt *p);

/* integrator: 0 leapfrog, 1 OMF4.  Output: hini, hfin, accepted (0/1). */
int orc_hmc_pure(double *U, double eps, int ns, int integrator, int qzero, uint64_t seed,
                 uint32_t chain, uint32_t traj, const orc_latt *p, double *hini,
                 double *hfin, int *accepted);
int orc_hmc_nf2(double *U, double am0, double eps, int ns, int integrator, int maxiter,
                double tol, int qzero, uint64_t seed, uint32_t chain, uint32_t traj,
                const orc_latt *p, double *hini, double *hfin, int *accepted, int *cg_iters);
/* dH of the last orc_hmc_nf2 trajectory from extended-precision energies. */
double orc_last_dh(void);
/* OpenMP threads a parallel region of the oracle uses (the CPU baseline's
 * thread count). */
int orc_omp_threads(void);
int orc_hmc_rhmc(double *U, const double *am0, int nf, const int *npole, double *const *mu,
                 double *const *nu, double *const *rho, double eps, int ns, int integrator,
                 int maxiter, double tol, int qzero, int reversibility_test, uint64_t seed,
                 uint32_t chain, uint32_t traj, const orc_latt *p, double *hini, double *hfin,
                 int *accepted);

/* ---- reweighting (src/QED2dRHMC.jl:174-251) ------------------------------ */
int orc_luscher_z(double *Z, const double *U, const double *X, double am0, int n,
                  const double *mu, const double *rho, double A, double r_b, int maxiter,
                  double tol, int L1, int L2, int bc);
int orc_luscher_zp(double *Z, int pw, const double *U, const double *X, double am0, int n,
                   const double *mu, const double *rho, double A, double r_b, int maxiter,
                   double tol, int L1, int L2, int bc);
double orc_reweighting_factor(const double *U, double am0, int n, const double *mu,
                              const double *rho, double A, double r_b, int N, int taylor,
                              int maxiter, double tol, uint64_t seed, uint32_t chain,
                              uint32_t traj, int L1, int L2, int bc, int *status);

/* ---- lattice I/O and master-field growth -------------------------------- */
void orc_unfold(double *Uout, const double *Uin, int L1, int L2, int reflect);

#ifdef __cplusplus
}
#endif
#endif
