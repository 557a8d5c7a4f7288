/*
 * qed2d_oracle.c -- FP64 CPU restatement of the reference QED2d MD path.
 *
 * TEST INFRASTRUCTURE ONLY (see qed2d_oracle.h for who may load it and what
 * pins each function).  Reference citations are to dalbandea/Dav-QED2d.
 *
 * Build: oracle/Makefile (gcc, -O2 -fopenmp -ffp-contract=off
 * -fcx-limited-range).  Site loops are OpenMP-parallel; every global sum
 * uses a fixed chunking with Neumaier compensation, so results do not depend
 * on the thread count.
 */
#include "qed2d_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Loops shorter than this run on one thread: a parallel region costs more
 * than it saves on a small lattice (24x24 ran 13x slower on 16 threads). */
#define OMP_MIN 8192

typedef double _Complex cplx;

#define IDX(i1, i2, c, L1, L2) ((((long)(c) * (L2) + (i2)) * (L1)) + (i1))

/* ========================================================================
 * Random numbers.  The reference draws from CURAND and Julia's Mersenne
 * Twister (src/QED2dHMC.jl:25,42,69-70,95); those streams are not
 * reproducible here, so this restatement defines counter-based streams:
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) + Box-Muller, with
 * log and sin/cos evaluated by fixed FMA sequences so that host and device
 * produce identical bits.
 * ======================================================================== */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words (exact in FP64). */
static double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1.0p-53;
}

/* log(u), u > 0: u = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(f),
 * f = (m-1)/(m+1), atanh series to f^23. */
double orc_det_log(double u)
{
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

/* sin(2 pi u), cos(2 pi u): 4u = n + r exactly, |r| <= 1/2, Taylor series of
 * sin(pi r/2) and cos(pi r/2) in r, then the quadrant n mod 4. */
void orc_det_sincos2pi(double u, double *s_out, double *c_out)
{
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
    switch (q) {
    case 0: *s_out = sv; *c_out = cv; break;
    case 1: *s_out = cv; *c_out = -sv; break;
    case 2: *s_out = -sv; *c_out = -cv; break;
    default: *s_out = -cv; *c_out = sv; break;
    }
}

static void make_key(uint64_t seed, uint32_t chain, uint32_t key[2])
{
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32) ^ chain;
}

void orc_gauss_pair(const uint32_t ctr[4], const uint32_t key[2], double *g1, double *g2)
{
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    double u1 = u53(x[0], x[1]);
    double u2 = u53(x[2], x[3]);
    double rad = sqrt(-2.0 * orc_det_log(1.0 - u1));
    double s, c;
    orc_det_sincos2pi(u2, &s, &c);
    *g1 = rad * c;
    *g2 = rad * s;
}

double orc_uniform(const uint32_t ctr[4], const uint32_t key[2])
{
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    return u53(x[0], x[1]);
}

enum { PURPOSE_MOM = 0x10, PURPOSE_PF = 0x20, PURPOSE_METRO = 0x30, PURPOSE_HOT = 0x40 };

/* mom = randn(L1,L2,2) (src/QED2dHMC.jl:25,69): one Philox block per site,
 * the two Gaussians go to directions 1 and 2. */
void orc_fill_momenta(double *mom, int L1, int L2, uint64_t seed, uint32_t chain, uint32_t traj)
{
    long V = (long)L1 * L2;
    uint32_t key[2];
    make_key(seed, chain, key);
#pragma omp parallel for if (V >= OMP_MIN)
    for (long s = 0; s < V; ++s) {
        uint32_t ctr[4] = {(uint32_t)s, 0u, PURPOSE_MOM, traj};
        orc_gauss_pair(ctr, key, &mom[s], &mom[V + s]);
    }
}

/* X = (randn + i randn)/sqrt(2) with the stream of one purpose: one Philox
 * block per site and spin component. */
void orc_fill_cgauss(double *X, int L1, int L2, uint64_t seed, uint32_t chain, uint32_t traj,
                     uint32_t purpose, uint32_t sub0)
{
    long V = (long)L1 * L2;
    uint32_t key[2];
    make_key(seed, chain, key);
    const double sq2 = sqrt(2.0);
#pragma omp parallel for if (V >= OMP_MIN)
    for (long s = 0; s < V; ++s) {
        for (uint32_t c = 0; c < 2; ++c) {
            uint32_t ctr[4] = {(uint32_t)s, sub0 + c, purpose, traj};
            double g1, g2;
            orc_gauss_pair(ctr, key, &g1, &g2);
            X[2 * (c * V + s)] = g1 / sq2;
            X[2 * (c * V + s) + 1] = g2 / sq2;
        }
    }
}

/* Pseudofermion noise (src/QED2dHMC.jl:70,138). */
void orc_fill_pseudofermion(double *X, int L1, int L2, uint64_t seed, uint32_t chain,
                            uint32_t traj, uint32_t flavour)
{
    orc_fill_cgauss(X, L1, L2, seed, chain, traj, PURPOSE_PF + flavour, 0);
}

/* r = rand() of the Metropolis step (src/QED2dHMC.jl:42,95,177). */
double orc_metropolis_uniform(uint64_t seed, uint32_t chain, uint32_t traj)
{
    uint32_t key[2];
    make_key(seed, chain, key);
    uint32_t ctr[4] = {0u, 0u, PURPOSE_METRO, traj};
    return orc_uniform(ctr, key);
}

/* Hot start: link angles uniform in [-pi, pi). */
void orc_hot_start(double *U, int L1, int L2, uint64_t seed, uint32_t chain)
{
    long V = (long)L1 * L2;
    uint32_t key[2];
    make_key(seed, chain, key);
#pragma omp parallel for if (V >= OMP_MIN)
    for (long s = 0; s < V; ++s) {
        for (uint32_t d = 0; d < 2; ++d) {
            uint32_t ctr[4] = {(uint32_t)s, d, PURPOSE_HOT, 0u};
            double u = orc_uniform(ctr, key);
            double sn, cs;
            orc_det_sincos2pi(u - 0.5, &sn, &cs);
            U[2 * (d * V + s)] = cs;
            U[2 * (d * V + s) + 1] = sn;
        }
    }
}

/* ========================================================================
 * Reductions: fixed 256-way chunking, Neumaier compensation inside a chunk,
 * chunks combined in order.  Independent of the OpenMP thread count.
 * ======================================================================== */
#define NCHUNK 256

static void neumaier_add(double *sum, double *comp, double x)
{
    double t = *sum + x;
    if (fabs(*sum) >= fabs(x)) *comp += (*sum - t) + x;
    else *comp += (x - t) + *sum;
    *sum = t;
}

typedef double (*term_fn)(const void *ctx, long i);

static double reduce_terms(term_fn f, const void *ctx, long n)
{
    double part[NCHUNK], cpart[NCHUNK];
#pragma omp parallel for if (n >= OMP_MIN)
    for (int ch = 0; ch < NCHUNK; ++ch) {
        long lo = n * ch / NCHUNK, hi = n * (ch + 1) / NCHUNK;
        double s = 0.0, c = 0.0;
        for (long i = lo; i < hi; ++i) neumaier_add(&s, &c, f(ctx, i));
        part[ch] = s;
        cpart[ch] = c;
    }
    double s = 0.0, c = 0.0;
    for (int ch = 0; ch < NCHUNK; ++ch) {
        neumaier_add(&s, &c, part[ch]);
        neumaier_add(&s, &c, cpart[ch]);
    }
    return s + c;
}

/* reduce_terms with its Neumaier sum and compensation added in long double
 * (x86-64: 64-bit mantissa), for energies whose differences must be resolved
 * below the ulp of a double. */
static long double reduce_terms_ld(term_fn f, const void *ctx, long n)
{
    double part[NCHUNK], cpart[NCHUNK];
#pragma omp parallel for if (n >= OMP_MIN)
    for (int ch = 0; ch < NCHUNK; ++ch) {
        long lo = n * ch / NCHUNK, hi = n * (ch + 1) / NCHUNK;
        double s = 0.0, c = 0.0;
        for (long i = lo; i < hi; ++i) neumaier_add(&s, &c, f(ctx, i));
        part[ch] = s;
        cpart[ch] = c;
    }
    long double s = 0.0L;
    for (int ch = 0; ch < NCHUNK; ++ch) s += (long double)part[ch] + (long double)cpart[ch];
    return s;
}

typedef struct { const double *a, *b; } pair_ctx;

static double term_abs2(const void *ctx, long i)
{
    const pair_ctx *p = (const pair_ctx *)ctx;
    return p->a[2 * i] * p->a[2 * i] + p->a[2 * i + 1] * p->a[2 * i + 1];
}
static double term_dot_re(const void *ctx, long i)
{
    const pair_ctx *p = (const pair_ctx *)ctx;
    return p->a[2 * i] * p->b[2 * i] + p->a[2 * i + 1] * p->b[2 * i + 1];
}
static double term_dot_im(const void *ctx, long i)
{
    const pair_ctx *p = (const pair_ctx *)ctx;
    return p->a[2 * i] * p->b[2 * i + 1] - p->a[2 * i + 1] * p->b[2 * i];
}
static double term_sq(const void *ctx, long i)
{
    const pair_ctx *p = (const pair_ctx *)ctx;
    return p->a[i] * p->a[i];
}
static double term_val(const void *ctx, long i)
{
    const pair_ctx *p = (const pair_ctx *)ctx;
    return p->a[i];
}

/* CUDA.mapreduce(abs2, +, a) (src/QED2dSolver.jl:13). n = complex count. */
double orc_norm2(const double *a, long n)
{
    pair_ctx c = {a, a};
    return reduce_terms(term_abs2, &c, n);
}

/* CUDA.dot(a, b) = sum conj(a) b (src/QED2dSolver.jl:21). */
void orc_dot(const double *a, const double *b, long n, double *re, double *im)
{
    pair_ctx c = {a, b};
    *re = reduce_terms(term_dot_re, &c, n);
    *im = reduce_terms(term_dot_im, &c, n);
}

static cplx dotc(const cplx *a, const cplx *b, long n)
{
    double re, im;
    orc_dot((const double *)a, (const double *)b, n, &re, &im);
    return CMPLX(re, im);
}

/* ========================================================================
 * Dirac operator
 * ======================================================================== */

/* Q = gamma5 Dw at every site: src/QED2dDirac.jl:2-33.  bc bit 0 makes the
 * hop across i2 = L2 -> 1 antiperiodic (a sign on U2 at i2 = L2). */
void orc_gamm5Dw(double *so_, const double *U_, const double *si_, double am0,
                 int L1, int L2, int bc)
{
    cplx *so = (cplx *)so_;
    const cplx *U = (const cplx *)U_, *si = (const cplx *)si_;
#pragma omp parallel for collapse(2) if ((long)L1 * L2 >= OMP_MIN)
    for (int i2 = 0; i2 < L2; ++i2) {
        for (int i1 = 0; i1 < L1; ++i1) {
            int iu1 = (i1 + 1) % L1, iu2 = (i2 + 1) % L2;
            int id1 = (i1 - 1 + L1) % L1, id2 = (i2 - 1 + L2) % L2;
            cplx U1 = U[IDX(i1, i2, 0, L1, L2)];
            cplx U1d = U[IDX(id1, i2, 0, L1, L2)];
            cplx U2 = U[IDX(i1, i2, 1, L1, L2)];
            cplx U2d = U[IDX(i1, id2, 1, L1, L2)];
            if ((bc & ORC_BC_APERIODIC_2) && i2 == L2 - 1) U2 = -U2;
            if ((bc & ORC_BC_APERIODIC_2) && i2 == 0) U2d = -U2d;
            cplx sp11 = si[IDX(iu1, i2, 0, L1, L2)], sp12 = si[IDX(iu1, i2, 1, L1, L2)];
            cplx sm11 = si[IDX(id1, i2, 0, L1, L2)], sm12 = si[IDX(id1, i2, 1, L1, L2)];
            cplx sp21 = si[IDX(i1, iu2, 0, L1, L2)], sp22 = si[IDX(i1, iu2, 1, L1, L2)];
            cplx sm21 = si[IDX(i1, id2, 0, L1, L2)], sm22 = si[IDX(i1, id2, 1, L1, L2)];
            cplx A = 0.5 * (U1 * (sp11 - sp12) + conj(U1d) * (sm11 + sm12));
            cplx B = 0.5 * (U2 * (sp21 + CMPLX(-cimag(sp22), creal(sp22))) +
                            conj(U2d) * (sm21 + CMPLX(cimag(sm22), -creal(sm22))));
            cplx A2 = 0.5 * (U1 * (sp11 - sp12) - conj(U1d) * (sm11 + sm12));
            cplx B2 = 0.5 * (U2 * (-sp22 + CMPLX(-cimag(sp21), creal(sp21))) -
                             conj(U2d) * (sm22 + CMPLX(-cimag(sm21), creal(sm21))));
            so[IDX(i1, i2, 0, L1, L2)] = (2.0 + am0) * si[IDX(i1, i2, 0, L1, L2)] - A - B;
            so[IDX(i1, i2, 1, L1, L2)] = -(2.0 + am0) * si[IDX(i1, i2, 1, L1, L2)] - A2 - B2;
        }
    }
}

/* gamma5: src/QED2dDirac.jl:130-146. */
void orc_gamm5(double *so_, const double *si_, int L1, int L2)
{
    cplx *so = (cplx *)so_;
    const cplx *si = (const cplx *)si_;
    long V = (long)L1 * L2;
    for (long s = 0; s < V; ++s) {
        so[s] = si[s];
        so[V + s] = -si[V + s];
    }
}

/* Dw = gamma5 (gamma5 Dw): src/QED2dDirac.jl:157-169. */
void orc_Dw(double *so, const double *U, const double *si, double am0, int L1, int L2, int bc)
{
    long n = 2L * L1 * L2;
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
    orc_gamm5Dw(so, U, si, am0, L1, L2, bc);
    memcpy(tmp, so, sizeof(double) * 2 * n);
    orc_gamm5(so, tmp, L1, L2);
    free(tmp);
}

/* Q^2: src/QED2dDirac.jl:35-46 (tmp allocated per call, as the reference). */
void orc_gamm5Dw_sqr(double *so, const double *U, const double *si, double am0,
                     int L1, int L2, int bc)
{
    long n = 2L * L1 * L2;
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
    orc_gamm5Dw(so, U, si, am0, L1, L2, bc);
    memcpy(tmp, so, sizeof(double) * 2 * n);
    orc_gamm5Dw(so, U, tmp, am0, L1, L2, bc);
    free(tmp);
}

/* Q^2 with caller's tmp: src/QED2dDirac.jl:69-81 (the Nf=2 CG operator). */
void orc_gamm5Dw_sqr_msq(double *so, double *tmp, const double *U, const double *si,
                         double am0, int L1, int L2, int bc)
{
    long n = 2L * L1 * L2;
    orc_gamm5Dw(so, U, si, am0, L1, L2, bc);
    memcpy(tmp, so, sizeof(double) * 2 * n);
    orc_gamm5Dw(so, U, tmp, am0, L1, L2, bc);
}

/* Q^2 + mu^2: src/QED2dDirac.jl:83-95 (the RHMC operator). */
void orc_gamm5Dw_sqr_musq(double *so, double *tmp, const double *U, const double *si,
                          double am0, double mu, int L1, int L2, int bc)
{
    long n = 2L * L1 * L2;
    orc_gamm5Dw_sqr_msq(so, tmp, U, si, am0, L1, L2, bc);
    double m2 = mu * mu;
    for (long i = 0; i < 2 * n; ++i) so[i] = so[i] + m2 * si[i];
}

/* Q^4: src/QED2dDirac.jl:48-67. */
void orc_gamm5Dw_sqr_sqr(double *so, const double *U, const double *si, double am0,
                         int L1, int L2, int bc)
{
    long n = 2L * L1 * L2;
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
    orc_gamm5Dw(so, U, si, am0, L1, L2, bc);
    for (int k = 0; k < 3; ++k) {
        memcpy(tmp, so, sizeof(double) * 2 * n);
        orc_gamm5Dw(so, U, tmp, am0, L1, L2, bc);
    }
    free(tmp);
}

/* Q^4 + mu^2: src/QED2dDirac.jl:97-117. */
void orc_gamm5Dw_sqr_sqr_musq(double *so, double *tmp, const double *U, const double *si,
                              double am0, double mu, int L1, int L2, int bc)
{
    long n = 2L * L1 * L2;
    orc_gamm5Dw(so, U, si, am0, L1, L2, bc);
    for (int k = 0; k < 3; ++k) {
        memcpy(tmp, so, sizeof(double) * 2 * n);
        orc_gamm5Dw(so, U, tmp, am0, L1, L2, bc);
    }
    double m2 = mu * mu;
    for (long i = 0; i < 2 * n; ++i) so[i] = so[i] + m2 * si[i];
}

/* ========================================================================
 * Forces, updates, observables
 * ======================================================================== */

/* Fermion force Re(i z_mu): src/QED2dDirac.jl:172-199.  Y = Q X. */
void orc_tr_dQwdU(double *frc, const double *U_, const double *X_, const double *Y_,
                  int L1, int L2, int bc)
{
    const cplx *U = (const cplx *)U_, *X = (const cplx *)X_, *Y = (const cplx *)Y_;
    long V = (long)L1 * L2;
    const cplx Im1 = CMPLX(0.0, 1.0);
#pragma omp parallel for collapse(2) if (V >= OMP_MIN)
    for (int i2 = 0; i2 < L2; ++i2) {
        for (int i1 = 0; i1 < L1; ++i1) {
            int iu1 = (i1 + 1) % L1, iu2 = (i2 + 1) % L2;
            cplx U1 = U[IDX(i1, i2, 0, L1, L2)], U2 = U[IDX(i1, i2, 1, L1, L2)];
            if ((bc & ORC_BC_APERIODIC_2) && i2 == L2 - 1) U2 = -U2;
            cplx X1 = X[IDX(i1, i2, 0, L1, L2)], X2 = X[IDX(i1, i2, 1, L1, L2)];
            cplx Y1 = Y[IDX(i1, i2, 0, L1, L2)], Y2 = Y[IDX(i1, i2, 1, L1, L2)];
            cplx Xu11 = X[IDX(iu1, i2, 0, L1, L2)], Xu12 = X[IDX(iu1, i2, 1, L1, L2)];
            cplx Yu11 = Y[IDX(iu1, i2, 0, L1, L2)], Yu12 = Y[IDX(iu1, i2, 1, L1, L2)];
            cplx Xu21 = X[IDX(i1, iu2, 0, L1, L2)], Xu22 = X[IDX(i1, iu2, 1, L1, L2)];
            cplx Yu21 = Y[IDX(i1, iu2, 0, L1, L2)], Yu22 = Y[IDX(i1, iu2, 1, L1, L2)];
            cplx z1 = conj(U1) * (conj(Xu11) * (Y1 + Y2) - conj(Xu12) * (Y1 + Y2)) -
                      U1 * (conj(X1) * (Yu11 - Yu12) + conj(X2) * (Yu11 - Yu12));
            cplx z2 = conj(U2) * (conj(Xu21) * (Y1 - Im1 * Y2) - conj(Xu22) * (Im1 * Y1 + Y2)) -
                      U2 * (conj(X1) * (Yu21 + Im1 * Yu22) + conj(X2) * (Im1 * Yu21 - Yu22));
            long s = (long)i2 * L1 + i1;
            frc[s] = creal(Im1 * z1);
            frc[V + s] = creal(Im1 * z2);
        }
    }
}

static inline cplx plaq(const cplx *U, int i1, int i2, int L1, int L2)
{
    int iu1 = (i1 + 1) % L1, iu2 = (i2 + 1) % L2;
    return U[IDX(i1, i2, 0, L1, L2)] * U[IDX(iu1, i2, 1, L1, L2)] *
           conj(U[IDX(i1, iu2, 0, L1, L2)] * U[IDX(i1, i2, 1, L1, L2)]);
}

/* Gauge force, scatter form: src/QED2dAction.jl:34-53. */
void orc_krnl_force(double *frc1, double *frc2, const double *U_, double beta, int L1, int L2)
{
    const cplx *U = (const cplx *)U_;
    long V = (long)L1 * L2;
#pragma omp parallel for collapse(2) if (V >= OMP_MIN)
    for (int i2 = 0; i2 < L2; ++i2) {
        for (int i1 = 0; i1 < L1; ++i1) {
            int iu1 = (i1 + 1) % L1, iu2 = (i2 + 1) % L2;
            double v = beta * cimag(plaq(U, i1, i2, L1, L2));
            frc1[(long)i2 * L1 + i1] = -v;
            frc1[V + (long)i2 * L1 + i1] = v;
            frc2[V + (long)i2 * L1 + iu1] = -v;
            frc2[(long)iu2 * L1 + i1] = v;
        }
    }
}

/* U <- exp(i eps p) U: src/QED2dAction.jl:55-64. */
void orc_updt_fld(double *U_, const double *mom, double eps, int L1, int L2)
{
    cplx *U = (cplx *)U_;
    long n = 2L * L1 * L2;
#pragma omp parallel for if (n >= OMP_MIN)
    for (long i = 0; i < n; ++i) U[i] = CMPLX(cos(eps * mom[i]), sin(eps * mom[i])) * U[i];
}

/* Re P(x): src/QED2dAction.jl:2-16. */
void orc_plaquette_field(double *plx, const double *U_, int L1, int L2)
{
    const cplx *U = (const cplx *)U_;
#pragma omp parallel for collapse(2) if ((long)L1 * L2 >= OMP_MIN)
    for (int i2 = 0; i2 < L2; ++i2)
        for (int i1 = 0; i1 < L1; ++i1) plx[(long)i2 * L1 + i1] = creal(plaq(U, i1, i2, L1, L2));
}

/* arg P(x): src/QED2dAction.jl:18-32. */
void orc_qtop_field(double *plx, const double *U_, int L1, int L2)
{
    const cplx *U = (const cplx *)U_;
#pragma omp parallel for collapse(2) if ((long)L1 * L2 >= OMP_MIN)
    for (int i2 = 0; i2 < L2; ++i2)
        for (int i1 = 0; i1 < L1; ++i1) plx[(long)i2 * L1 + i1] = carg(plaq(U, i1, i2, L1, L2));
}

/* S = beta (V - sum Re P): src/QED2dHMC.jl:206-215. */
double orc_action(const double *U, double beta, int L1, int L2)
{
    long V = (long)L1 * L2;
    double *plx = (double *)malloc(sizeof(double) * V);
    orc_plaquette_field(plx, U, L1, L2);
    pair_ctx c = {plx, plx};
    double s = reduce_terms(term_val, &c, V);
    free(plx);
    return beta * ((double)V - s);
}

/* <P> = sum Re P / V: src/QED2dHMC.jl:217-226. */
double orc_plaquette(const double *U, int L1, int L2)
{
    long V = (long)L1 * L2;
    double *plx = (double *)malloc(sizeof(double) * V);
    orc_plaquette_field(plx, U, L1, L2);
    pair_ctx c = {plx, plx};
    double s = reduce_terms(term_val, &c, V);
    free(plx);
    return s / (double)V;
}

/* Q = sum arg P / 2 pi: src/QED2dHMC.jl:228-237. */
double orc_qtop(const double *U, int L1, int L2)
{
    long V = (long)L1 * L2;
    double *plx = (double *)malloc(sizeof(double) * V);
    orc_qtop_field(plx, U, L1, L2);
    pair_ctx c = {plx, plx};
    double s = reduce_terms(term_val, &c, V);
    free(plx);
    return s / (2.0 * M_PI);
}

/* H = sum p^2 / 2 + S: src/QED2dHMC.jl:198-203. */
double orc_hamiltonian(const double *mom, const double *U, double beta, int L1, int L2)
{
    long V = (long)L1 * L2;
    pair_ctx c = {mom, mom};
    return reduce_terms(term_sq, &c, 2 * V) / 2.0 + orc_action(U, beta, L1, L2);
}

/* ========================================================================
 * Solvers
 * ======================================================================== */

static void apply_op(int op, double *so, double *tmp, const double *U, const double *si,
                     double am0, double mu, int L1, int L2, int bc)
{
    if (op == 0) orc_gamm5Dw_sqr_msq(so, tmp, U, si, am0, L1, L2, bc);
    else orc_gamm5Dw_sqr_musq(so, tmp, U, si, am0, mu, L1, L2, bc);
}

/* CG: v1 src/QED2dSolver.jl:3-45 (so_as_guess = 0, op = 0) and v2 49-105.
 * alpha = norm / dot(p, Ap) is complex, beta = err / norm real; stop when
 * |r|^2 < eps |b|^2. */
int orc_cg(double *so_, const double *U, const double *si_, double am0, double mu, int op,
           int maxiter, double eps, int so_as_guess, int L1, int L2, int bc, int *status)
{
    long n = 2L * L1 * L2;
    cplx *so = (cplx *)so_;
    const cplx *si = (const cplx *)si_;
    cplx *r = (cplx *)malloc(sizeof(cplx) * n);
    cplx *p = (cplx *)malloc(sizeof(cplx) * n);
    cplx *Ap = (cplx *)malloc(sizeof(cplx) * n);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * n);
    double tol = eps * orc_norm2(si_, n);
    double norm;
    int iterations = 0;
    *status = 0;
    if (!so_as_guess) {
        for (long i = 0; i < n; ++i) { so[i] = 0.0; r[i] = si[i]; p[i] = si[i]; }
        norm = orc_norm2(si_, n);
    } else {
        apply_op(op, (double *)Ap, (double *)tmp, U, so_, am0, mu, L1, L2, bc);
        for (long i = 0; i < n; ++i) { r[i] = si[i] - Ap[i]; p[i] = r[i]; }
        norm = orc_norm2((double *)r, n);
        if (norm < tol) goto done;
    }
    double err = 0.0;
    for (int it = 1; it <= maxiter; ++it) {
        apply_op(op, (double *)Ap, (double *)tmp, U, (double *)p, am0, mu, L1, L2, bc);
        cplx prod = dotc(p, Ap, n);
        cplx alpha = norm / prod;
#pragma omp parallel for if (n >= OMP_MIN)
        for (long i = 0; i < n; ++i) {
            so[i] = so[i] + alpha * p[i];
            r[i] = r[i] - alpha * Ap[i];
        }
        err = orc_norm2((double *)r, n);
        if (err < tol) { iterations = it; break; }
        double beta = err / norm;
#pragma omp parallel for if (n >= OMP_MIN)
        for (long i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        norm = err;
    }
    if (err > tol) *status = 1;
done:
    free(r); free(p); free(Ap); free(tmp);
    return iterations;
}

/* MultiCG: so = (1 + sum_j rho_j (Q^2 + mu_j^2)^-1) si: src/QED2dRHMC.jl:281-290. */
int orc_multicg(double *so, const double *U, const double *si, double am0, int n,
                const double *mu, const double *rho, int maxiter, double eps,
                int L1, int L2, int bc)
{
    long m = 4L * L1 * L2;
    double *aux = (double *)malloc(sizeof(double) * m);
    int st = 0, bad = 0;
    memcpy(so, si, sizeof(double) * m);
    for (int j = 0; j < n; ++j) {
        orc_cg(aux, U, si, am0, mu[j], 1, maxiter, eps, 0, L1, L2, bc, &st);
        bad |= st;
        for (long i = 0; i < m; ++i) so[i] = so[i] + rho[j] * aux[i];
    }
    free(aux);
    return bad;
}

/* R = (A / r_b) MultiCG: src/QED2dRHMC.jl:261-271. */
int orc_R(double *so, const double *U, const double *si, double am0, int n,
          const double *mu, const double *rho, double A, double r_b, int maxiter,
          double eps, int L1, int L2, int bc)
{
    long m = 4L * L1 * L2;
    int bad = orc_multicg(so, U, si, am0, n, mu, rho, maxiter, eps, L1, L2, bc);
    double f = pow(r_b, -1.0) * A;
    for (long i = 0; i < m; ++i) so[i] = f * so[i];
    return bad;
}

/* phi = prod_i (Q - i nu_i)(Q^2 + nu_i^2)^-1 (Q + i mu_i) X:
 * src/QED2dRHMC.jl:298-319. */
int orc_generate_pseudofermion(double *F_, const double *U, const double *X, double am0,
                               int n, const double *mu, const double *nu, int maxiter,
                               double tol, int L1, int L2, int bc)
{
    long m = 2L * L1 * L2;
    cplx *F = (cplx *)F_;
    cplx *aux = (cplx *)malloc(sizeof(cplx) * m);
    int st = 0, bad = 0;
    memcpy(F, X, sizeof(cplx) * m);
    for (int i = 0; i < n; ++i) {
        memcpy(aux, F, sizeof(cplx) * m);
        orc_gamm5Dw(F_, U, (double *)aux, am0, L1, L2, bc);
        for (long k = 0; k < m; ++k) F[k] = F[k] + I * mu[i] * aux[k];
        memcpy(aux, F, sizeof(cplx) * m);
        orc_cg(F_, U, (double *)aux, am0, nu[i], 1, maxiter, tol, 0, L1, L2, bc, &st);
        bad |= st;
        memcpy(aux, F, sizeof(cplx) * m);
        orc_gamm5Dw(F_, U, (double *)aux, am0, L1, L2, bc);
        for (long k = 0; k < m; ++k) F[k] = F[k] - I * nu[i] * aux[k];
    }
    free(aux);
    return bad;
}

/* Spectral bounds of Q^2 by shifted power iteration: src/QED2dRHMC.jl:135-170.
 * b0 is the starting vector (the reference draws it at :137).  The second loop
 * applies Q^2 to a copy, where the reference aliases output and input (:159). */
void orc_power_method(const double *U, const double *b0, double am0, int iter, int L1,
                      int L2, int bc, double *lmin, double *lmax)
{
    long n = 2L * L1 * L2;
    cplx *b = (cplx *)malloc(sizeof(cplx) * n);
    cplx *ba = (cplx *)malloc(sizeof(cplx) * n);
    cplx *bl = (cplx *)malloc(sizeof(cplx) * n);
    memcpy(b, b0, sizeof(cplx) * n);
    const double shift = 2.0;
    for (int it = 0; it < iter; ++it) {
        orc_gamm5Dw_sqr((double *)ba, U, (double *)b, am0, L1, L2, bc);
        for (long i = 0; i < n; ++i) ba[i] = ba[i] + shift * b[i];
        cplx d = dotc(ba, ba, n);
        for (long i = 0; i < n; ++i) b[i] = ba[i] / d;
    }
    orc_gamm5Dw_sqr((double *)ba, U, (double *)b, am0, L1, L2, bc);
    for (long i = 0; i < n; ++i) ba[i] = ba[i] + shift * b[i];
    cplx lam_max = dotc(b, ba, n) / dotc(b, b, n) - shift;
    for (int it = 0; it < iter; ++it) {
        memcpy(bl, b, sizeof(cplx) * n);
        orc_gamm5Dw_sqr((double *)ba, U, (double *)bl, am0, L1, L2, bc);
        for (long i = 0; i < n; ++i) b[i] = ba[i] - lam_max * bl[i];
        cplx d = dotc(b, b, n);
        for (long i = 0; i < n; ++i) b[i] = b[i] / d;
    }
    orc_gamm5Dw_sqr((double *)ba, U, (double *)b, am0, L1, L2, bc);
    for (long i = 0; i < n; ++i) ba[i] = ba[i] - lam_max * b[i];
    cplx lam_min = dotc(b, ba, n) / dotc(b, b, n) + lam_max;
    *lmin = creal(lam_min);
    *lmax = creal(lam_max);
    free(b); free(ba); free(bl);
}

/* ========================================================================
 * Zolotarev rational approximation of y^(-1/2): src/QED2dRHMC.jl:2-121.
 * The reference takes K, sn, cn from Elliptic.jl 1.0.1 (parameter
 * convention m = k^2).  Restated here from the arithmetic-geometric mean
 * (Abramowitz & Stegun 17.6) and the descending Landen / AGM scheme for the
 * Jacobi functions (A&S 16.4).
 * ======================================================================== */
double orc_ellipk(double m)
{
    double a = 1.0, b = sqrt(1.0 - m);
    for (int i = 0; i < 64 && fabs(a - b) > 1e-17 * a; ++i) {
        double an = 0.5 * (a + b);
        b = sqrt(a * b);
        a = an;
    }
    return M_PI / (2.0 * a);
}

void orc_ellip_sncn(double u, double m, double *sn, double *cn)
{
    double a[80], c[80];
    double b = sqrt(1.0 - m);
    a[0] = 1.0;
    c[0] = sqrt(m);
    int N = 0;
    while (N < 70 && fabs(c[N]) > 1e-17 * a[N]) {
        a[N + 1] = 0.5 * (a[N] + b);
        c[N + 1] = 0.5 * (a[N] - b);
        b = sqrt(a[N] * b);
        ++N;
    }
    double phi = ldexp(a[N] * u, N);
    for (int k = N; k >= 1; --k) phi = 0.5 * (phi + asin(c[k] / a[k] * sin(phi)));
    *sn = sin(phi);
    *cn = cos(phi);
}

typedef struct { double eps, k, v, m; int n; } zolo;

static double zk(double eps) { return sqrt(1.0 - eps); }

static void zolo_init(zolo *z, int n, double eps)
{
    z->n = n;
    z->eps = eps;
    z->k = zk(eps);
    z->m = z->k * z->k;
    z->v = orc_ellipk(z->m) / (2.0 * n + 1.0);
}

static double z_a_r(const zolo *z, int r)
{
    double s, c;
    orc_ellip_sncn(r * z->v, z->m, &s, &c);
    return (c * c) / (s * s);
}
static double z_c_r(const zolo *z, int r)
{
    double s, c;
    orc_ellip_sncn(r * z->v, z->m, &s, &c);
    return s * s;
}
static double z_mu(const zolo *z, int j, double r_b) { return r_b * pow(z_a_r(z, 2 * j), 0.5); }
static double z_nu(const zolo *z, int j, double r_b) { return r_b * pow(z_a_r(z, 2 * j - 1), 0.5); }

static double z_rho_mu(const zolo *z, int j, int kk, int l, double r_b)
{
    double muj = z_mu(z, j, r_b), nuj = z_nu(z, j, r_b);
    double res = nuj * nuj - muj * muj;
    for (int m = kk; m <= l; ++m) {
        if (m == j) continue;
        double num = z_nu(z, m, r_b), mum = z_mu(z, m, r_b);
        res *= (num * num - muj * muj) / (mum * mum - muj * muj);
    }
    return res;
}

static double z_d(const zolo *z)
{
    double res = pow(z->k, 2 * z->n + 1);
    for (int i = 1; i <= 2 * z->n + 1; i += 2) {
        double c = z_c_r(z, i);
        res *= c * c;
    }
    return res;
}

void orc_zolotarev(int n, double r_a, double r_b, double *eps_out, double *A_out,
                   double *delta_out, double *rho, double *mu, double *nu)
{
    double eps = (r_a / r_b) * (r_a / r_b);
    zolo z;
    zolo_init(&z, n, eps);
    double d = z_d(&z);
    double delta = d * d / ((1.0 + sqrt(1.0 - d * d)) * (1.0 + sqrt(1.0 - d * d)));
    double A = 2.0 / (1.0 + sqrt(1.0 - d * d));
    for (int i = 1; i <= 2 * n - 1; i += 2) A *= z_c_r(&z, i);
    for (int i = 2; i <= 2 * n; i += 2) A *= 1.0 / z_c_r(&z, i);
    for (int j = 1; j <= n; ++j) {
        mu[j - 1] = z_mu(&z, j, r_b);
        nu[j - 1] = z_nu(&z, j, r_b);
        rho[j - 1] = z_rho_mu(&z, j, 1, n, r_b);
    }
    *eps_out = eps;
    *A_out = A;
    *delta_out = delta;
}

/* ========================================================================
 * Molecular dynamics
 * ======================================================================== */

static const double OMF_R1 = 0.08398315262876693;
static const double OMF_R2 = 0.2539785108410595;
static const double OMF_R3 = 0.6822365335719091;
static const double OMF_R4 = -0.03230286765269967;

/* Nf=2 force + momentum update: src/QED2dHMC.jl:452-481. */
int orc_update_momenta_nf2(double *mom, const double *U, double *X, const double *F,
                           double *g5DX, double am0, double eps, int maxiter, double tol,
                           const orc_latt *p)
{
    long V = (long)p->L1 * p->L2;
    double *frc1 = (double *)calloc(2 * V, sizeof(double));
    double *frc2 = (double *)calloc(2 * V, sizeof(double));
    double *frc = (double *)calloc(2 * V, sizeof(double));
    int st = 0;
    int it = orc_cg(X, U, F, am0, 0.0, 0, maxiter, tol, 0, p->L1, p->L2, p->bc, &st);
    orc_gamm5Dw(g5DX, U, X, am0, p->L1, p->L2, p->bc);
    orc_tr_dQwdU(frc, U, X, g5DX, p->L1, p->L2, p->bc);
    orc_krnl_force(frc1, frc2, U, p->beta, p->L1, p->L2);
#pragma omp parallel for if (V >= OMP_MIN)
    for (long i = 0; i < 2 * V; ++i) mom[i] = mom[i] + eps * (frc1[i] + frc2[i] + frc[i]);
    free(frc1); free(frc2); free(frc);
    return st ? -1 : it;
}

/* Leapfrog for Nf=2: src/QED2dHMC.jl:426-449.  Returns total CG iterations
 * (negative on non-convergence). */
int orc_leapfrog_nf2(double *mom, double *U, double *X, const double *F, double *g5DX,
                     double am0, double eps, int ns, int maxiter, double tol,
                     const orc_latt *p)
{
    int tot = 0, it;
#define ACC(x) do { it = (x); if (it < 0) return -1; tot += it; } while (0)
    ACC(orc_update_momenta_nf2(mom, U, X, F, g5DX, am0, 0.5 * eps, maxiter, tol, p));
    for (int i = 0; i < ns - 1; ++i) {
        orc_updt_fld(U, mom, eps, p->L1, p->L2);
        ACC(orc_update_momenta_nf2(mom, U, X, F, g5DX, am0, eps, maxiter, tol, p));
    }
    orc_updt_fld(U, mom, eps, p->L1, p->L2);
    ACC(orc_update_momenta_nf2(mom, U, X, F, g5DX, am0, 0.5 * eps, maxiter, tol, p));
    return tot;
}

/* OMF4 for Nf=2: src/QED2dHMC.jl:307-422. */
int orc_omf4_nf2(double *mom, double *U, double *X, const double *F, double *g5DX,
                 double am0, double eps, int ns, int maxiter, double tol, const orc_latt *p)
{
    const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
    const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
    const double pc[6] = {r1, r3, r5, r5, r3, r1};
    const double uc[5] = {r2, r4, r6, r4, r2};
    int tot = 0, it;
    for (int i = 0; i < ns; ++i) {
        for (int st = 0; st < 6; ++st) {
            ACC(orc_update_momenta_nf2(mom, U, X, F, g5DX, am0, pc[st] * eps, maxiter, tol, p));
            if (st < 5) orc_updt_fld(U, mom, eps * uc[st], p->L1, p->L2);
        }
    }
    return tot;
}

static void pure_momenta(double *mom, const double *U, double *frc1, double *frc2,
                         double coef, const orc_latt *p)
{
    long V = (long)p->L1 * p->L2;
    orc_krnl_force(frc1, frc2, U, p->beta, p->L1, p->L2);
#pragma omp parallel for if (V >= OMP_MIN)
    for (long i = 0; i < 2 * V; ++i) mom[i] = mom[i] + coef * (frc1[i] + frc2[i]);
}

/* OMF4 for the pure gauge theory: src/QED2dHMC.jl:239-305. */
void orc_omf4_pure(double *mom, double *U, double eps, int ns, const orc_latt *p)
{
    long V = (long)p->L1 * p->L2;
    const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
    const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
    const double pc[6] = {r1, r3, r5, r5, r3, r1};
    const double uc[5] = {r2, r4, r6, r4, r2};
    double *frc1 = (double *)calloc(2 * V, sizeof(double));
    double *frc2 = (double *)calloc(2 * V, sizeof(double));
    for (int i = 0; i < ns; ++i) {
        for (int st = 0; st < 6; ++st) {
            pure_momenta(mom, U, frc1, frc2, pc[st] * eps, p);
            if (st < 5) orc_updt_fld(U, mom, eps * uc[st], p->L1, p->L2);
        }
    }
    free(frc1); free(frc2);
}

/* Leapfrog for the pure gauge theory (the reference's MD! dispatch has no
 * pure-gauge leapfrog; this follows the Nf=2 leapfrog stage order of
 * src/QED2dHMC.jl:426-449 with the gauge force only). */
void orc_leapfrog_pure(double *mom, double *U, double eps, int ns, const orc_latt *p)
{
    long V = (long)p->L1 * p->L2;
    double *frc1 = (double *)calloc(2 * V, sizeof(double));
    double *frc2 = (double *)calloc(2 * V, sizeof(double));
    pure_momenta(mom, U, frc1, frc2, 0.5 * eps, p);
    for (int i = 0; i < ns - 1; ++i) {
        orc_updt_fld(U, mom, eps, p->L1, p->L2);
        pure_momenta(mom, U, frc1, frc2, eps, p);
    }
    orc_updt_fld(U, mom, eps, p->L1, p->L2);
    pure_momenta(mom, U, frc1, frc2, 0.5 * eps, p);
    free(frc1); free(frc2);
}

/* RHMC force + momentum update: src/QED2dHMC.jl:516-557.  Warm start from the
 * previous pole's solution within a flavour (:526-545). */
int orc_update_momenta_rhmc(double *mom, const double *U, double *X, double *const *F,
                            double *g5DX, const double *am0, int nf, const int *npole,
                            double *const *mu, double *const *rho, double eps, int maxiter,
                            double tol, const orc_latt *p)
{
    long V = (long)p->L1 * p->L2;
    double *frc1 = (double *)calloc(2 * V, sizeof(double));
    double *frc2 = (double *)calloc(2 * V, sizeof(double));
    double *frci = (double *)calloc(2 * V, sizeof(double));
    double *frc = (double *)calloc(2 * V, sizeof(double));
    int iter = 0, st = 0, bad = 0;
    int guess = 0;
    for (int j = 0; j < nf; ++j) {
        for (int i = 0; i < npole[j]; ++i) {
            iter += orc_cg(X, U, F[j], am0[j], mu[j][i], 1, maxiter, tol, guess, p->L1, p->L2,
                           p->bc, &st);
            bad |= st;
            guess = 1;
            orc_gamm5Dw(g5DX, U, X, am0[j], p->L1, p->L2, p->bc);
            orc_tr_dQwdU(frci, U, X, g5DX, p->L1, p->L2, p->bc);
            for (long k = 0; k < 2 * V; ++k) frc[k] = frc[k] + rho[j][i] * frci[k];
        }
        guess = 0;
    }
    orc_krnl_force(frc1, frc2, U, p->beta, p->L1, p->L2);
    for (long k = 0; k < 2 * V; ++k) mom[k] = mom[k] + eps * (frc1[k] + frc2[k] + frc[k]);
    free(frc1); free(frc2); free(frci); free(frc);
    return bad ? -1 : iter;
}

/* Leapfrog for RHMC: src/QED2dHMC.jl:485-513. */
int orc_leapfrog_rhmc(double *mom, double *U, double *X, double *const *F, double *g5DX,
                      const double *am0, int nf, const int *npole, double *const *mu,
                      double *const *rho, double eps, int ns, int maxiter, double tol,
                      const orc_latt *p)
{
    int tot = 0, it;
    ACC(orc_update_momenta_rhmc(mom, U, X, F, g5DX, am0, nf, npole, mu, rho, 0.5 * eps,
                                maxiter, tol, p));
    for (int i = 0; i < ns - 1; ++i) {
        orc_updt_fld(U, mom, eps, p->L1, p->L2);
        ACC(orc_update_momenta_rhmc(mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, maxiter,
                                    tol, p));
    }
    orc_updt_fld(U, mom, eps, p->L1, p->L2);
    ACC(orc_update_momenta_rhmc(mom, U, X, F, g5DX, am0, nf, npole, mu, rho, 0.5 * eps,
                                maxiter, tol, p));
    return tot;
#undef ACC
}

/* OMF4 with the RHMC force.  The reference has no such method (MD! dispatch,
 * src/QED2dHMC.jl:12-16); this is a build extension combining the OMF4 stage
 * order of src/QED2dHMC.jl:307-422 with the RHMC force of :516-557. */
int orc_omf4_rhmc(double *mom, double *U, double *X, double *const *F, double *g5DX,
                  const double *am0, int nf, const int *npole, double *const *mu,
                  double *const *rho, double eps, int ns, int maxiter, double tol,
                  const orc_latt *p)
{
    const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
    const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
    const double pc[6] = {r1, r3, r5, r5, r3, r1};
    const double uc[5] = {r2, r4, r6, r4, r2};
    int tot = 0;
    for (int i = 0; i < ns; ++i) {
        for (int st = 0; st < 6; ++st) {
            int it = orc_update_momenta_rhmc(mom, U, X, F, g5DX, am0, nf, npole, mu, rho,
                                             pc[st] * eps, maxiter, tol, p);
            if (it < 0) return -1;
            tot += it;
            if (st < 5) orc_updt_fld(U, mom, eps * uc[st], p->L1, p->L2);
        }
    }
    return tot;
}

/* Metropolis step: src/QED2dHMC.jl:31-51 (same for :84-104, :166-186). */
static int metropolis(double hini, double hfin, int qzero, const double *U, const orc_latt *p,
                      uint64_t seed, uint32_t chain, uint32_t traj)
{
    double pacc;
    if (qzero && fabs(orc_qtop(U, p->L1, p->L2)) > 0.1) pacc = 0.0;
    else pacc = exp(-(hfin - hini));
    if (pacc < 1.0) {
        double r = orc_metropolis_uniform(seed, chain, traj);
        return pacc > r ? 1 : 0;
    }
    return 1;
}

/* Pure-gauge HMC step: src/QED2dHMC.jl:19-60. */
int orc_hmc_pure(double *U, double eps, int ns, int integrator, int qzero, uint64_t seed,
                 uint32_t chain, uint32_t traj, const orc_latt *p, double *hini,
                 double *hfin, int *accepted)
{
    long V = (long)p->L1 * p->L2;
    double *Ucp = (double *)malloc(sizeof(double) * 4 * V);
    double *mom = (double *)malloc(sizeof(double) * 2 * V);
    memcpy(Ucp, U, sizeof(double) * 4 * V);
    orc_fill_momenta(mom, p->L1, p->L2, seed, chain, traj);
    *hini = orc_hamiltonian(mom, U, p->beta, p->L1, p->L2);
    if (integrator == 1) orc_omf4_pure(mom, U, eps, ns, p);
    else orc_leapfrog_pure(mom, U, eps, ns, p);
    *hfin = orc_hamiltonian(mom, U, p->beta, p->L1, p->L2);
    *accepted = metropolis(*hini, *hfin, qzero, U, p, seed, chain, traj);
    if (!*accepted) memcpy(U, Ucp, sizeof(double) * 4 * V);
    free(Ucp); free(mom);
    return 0;
}

/* H + Re X^dag Y (src/QED2dHMC.jl:74,82) in extended precision: the
 * compensated sums of orc_hamiltonian and orc_dot combined in long double, so
 * that the difference of two such energies of order 1e6 keeps ~1e-12. */
static long double hamiltonian_f_ld(const double *mom, const double *U, double beta, int L1,
                                    int L2, const double *X, const double *Y)
{
    long V = (long)L1 * L2;
    pair_ctx m = {mom, mom};
    long double kin = reduce_terms_ld(term_sq, &m, 2 * V) / 2.0L;
    double *plx = (double *)malloc(sizeof(double) * V);
    orc_plaquette_field(plx, U, L1, L2);
    pair_ctx pc = {plx, plx};
    long double sp = reduce_terms_ld(term_val, &pc, V);
    free(plx);
    pair_ctx f = {X, Y};
    long double xf = reduce_terms_ld(term_dot_re, &f, 2 * V);
    return kin + (long double)beta * ((long double)V - sp) + xf;
}

static double last_dh;

/* dH of the last orc_hmc_nf2 trajectory from the extended-precision energies
 * (hfin - hini of the returned doubles loses the bits below the ulp of H). */
double orc_last_dh(void) { return last_dh; }

/* Nf=2 HMC step: src/QED2dHMC.jl:64-113.  hfin uses the last MD solve (:82). */
int orc_hmc_nf2(double *U, double am0, double eps, int ns, int integrator, int maxiter,
                double tol, int qzero, uint64_t seed, uint32_t chain, uint32_t traj,
                const orc_latt *p, double *hini, double *hfin, int *accepted, int *cg_iters)
{
    long V = (long)p->L1 * p->L2;
    double *Ucp = (double *)malloc(sizeof(double) * 4 * V);
    double *mom = (double *)malloc(sizeof(double) * 2 * V);
    double *X = (double *)malloc(sizeof(double) * 4 * V);
    double *F = (double *)calloc(4 * V, sizeof(double));
    double *g5DX = (double *)calloc(4 * V, sizeof(double));
    memcpy(Ucp, U, sizeof(double) * 4 * V);
    orc_fill_momenta(mom, p->L1, p->L2, seed, chain, traj);
    orc_fill_pseudofermion(X, p->L1, p->L2, seed, chain, traj, 0);
    double xr, xi, fr, fi;
    orc_dot(X, X, 2 * V, &xr, &xi);
    *hini = orc_hamiltonian(mom, U, p->beta, p->L1, p->L2) + xr;
    const long double hini_ld = hamiltonian_f_ld(mom, U, p->beta, p->L1, p->L2, X, X);
    last_dh = 0.0;
    orc_gamm5Dw(F, U, X, am0, p->L1, p->L2, p->bc);
    int it;
    if (integrator == 1) it = orc_omf4_nf2(mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, p);
    else it = orc_leapfrog_nf2(mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, p);
    if (cg_iters) *cg_iters = it;
    int rc = 0;
    if (it < 0) { rc = 1; memcpy(U, Ucp, sizeof(double) * 4 * V); *accepted = 0; goto out; }
    orc_dot(X, F, 2 * V, &fr, &fi);
    *hfin = orc_hamiltonian(mom, U, p->beta, p->L1, p->L2) + fr;
    last_dh = (double)(hamiltonian_f_ld(mom, U, p->beta, p->L1, p->L2, X, F) - hini_ld);
    *accepted = metropolis(*hini, *hfin, qzero, U, p, seed, chain, traj);
    if (!*accepted) memcpy(U, Ucp, sizeof(double) * 4 * V);
out:
    free(Ucp); free(mom); free(X); free(F); free(g5DX);
    return rc;
}

/* RHMC step, one pseudofermion per flavour: src/QED2dHMC.jl:116-195. */
static int md_rhmc(int integrator, double *mom, double *U, double *X, double *const *F,
                   double *g5DX, const double *am0, int nf, const int *npole, double *const *mu,
                   double *const *rho, double eps, int ns, int maxiter, double tol,
                   const orc_latt *p)
{
    if (integrator == 1)
        return orc_omf4_rhmc(mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, ns, maxiter, tol, p);
    return orc_leapfrog_rhmc(mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, ns, maxiter, tol, p);
}

int orc_hmc_rhmc(double *U, const double *am0, int nf, const int *npole, double *const *mu,
                 double *const *nu, double *const *rho, double eps, int ns, int integrator,
                 int maxiter, double tol, int qzero, int reversibility_test, uint64_t seed,
                 uint32_t chain, uint32_t traj, const orc_latt *p, double *hini, double *hfin,
                 int *accepted)
{
    long V = (long)p->L1 * p->L2;
    double *Ucp = (double *)malloc(sizeof(double) * 4 * V);
    double *mom = (double *)malloc(sizeof(double) * 2 * V);
    double *X = (double *)calloc(4 * V, sizeof(double));
    double *g5DX = (double *)calloc(4 * V, sizeof(double));
    double **F = (double **)malloc(sizeof(double *) * nf);
    int bad = 0;
    memcpy(Ucp, U, sizeof(double) * 4 * V);
    orc_fill_momenta(mom, p->L1, p->L2, seed, chain, traj);
    *hini = orc_hamiltonian(mom, U, p->beta, p->L1, p->L2);
    for (int j = 0; j < nf; ++j) {
        orc_fill_pseudofermion(X, p->L1, p->L2, seed, chain, traj, (uint32_t)j);
        *hini += orc_norm2(X, 2 * V);
        F[j] = (double *)calloc(4 * V, sizeof(double));
        bad |= orc_generate_pseudofermion(F[j], U, X, am0[j], npole[j], mu[j], nu[j], maxiter,
                                          tol, p->L1, p->L2, p->bc);
    }
    if (md_rhmc(integrator, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, ns, maxiter, tol, p) < 0)
        bad = 1;
    *hfin = orc_hamiltonian(mom, U, p->beta, p->L1, p->L2);
    for (int j = 0; j < nf; ++j) {
        bad |= orc_multicg(X, U, F[j], am0[j], npole[j], mu[j], rho[j], maxiter, tol, p->L1,
                           p->L2, p->bc);
        double re, im;
        orc_dot(X, F[j], 2 * V, &re, &im);
        *hfin += re;
    }
    if (reversibility_test) {
        for (long k = 0; k < 2 * V; ++k) mom[k] = -mom[k];
        if (md_rhmc(integrator, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, ns, maxiter, tol,
                    p) < 0)
            bad = 1;
        *accepted = -1;
    } else {
        *accepted = metropolis(*hini, *hfin, qzero, U, p, seed, chain, traj);
        if (!*accepted) memcpy(U, Ucp, sizeof(double) * 4 * V);
    }
    for (int j = 0; j < nf; ++j) free(F[j]);
    free(F); free(Ucp); free(mom); free(X); free(g5DX);
    return bad;
}

/* Z X = (Q^2 R^2 - 1) X: src/QED2dRHMC.jl:174-186. */
int orc_luscher_z(double *Z, const double *U, const double *X, double am0, int n,
                  const double *mu, const double *rho, double A, double r_b, int maxiter,
                  double tol, int L1, int L2, int bc)
{
    long m = 4L * L1 * L2;
    double *tmp = (double *)malloc(sizeof(double) * m);
    int bad = orc_R(Z, U, X, am0, n, mu, rho, A, r_b, maxiter, tol, L1, L2, bc);
    memcpy(tmp, Z, sizeof(double) * m);
    bad |= orc_R(Z, U, tmp, am0, n, mu, rho, A, r_b, maxiter, tol, L1, L2, bc);
    memcpy(tmp, Z, sizeof(double) * m);
    orc_gamm5Dw(Z, U, tmp, am0, L1, L2, bc);
    memcpy(tmp, Z, sizeof(double) * m);
    orc_gamm5Dw(Z, U, tmp, am0, L1, L2, bc);
    for (long i = 0; i < m; ++i) Z[i] = Z[i] - X[i];
    free(tmp);
    return bad;
}

/* Z^p X: src/QED2dRHMC.jl:189-201. */
int orc_luscher_zp(double *Z, int pw, const double *U, const double *X, double am0, int n,
                   const double *mu, const double *rho, double A, double r_b, int maxiter,
                   double tol, int L1, int L2, int bc)
{
    long m = 4L * L1 * L2;
    double *tmp = (double *)malloc(sizeof(double) * m);
    int bad = 0;
    memcpy(tmp, X, sizeof(double) * m);
    for (int i = 0; i < pw; ++i) {
        bad |= orc_luscher_z(Z, U, tmp, am0, n, mu, rho, A, r_b, maxiter, tol, L1, L2, bc);
        memcpy(tmp, Z, sizeof(double) * m);
    }
    free(tmp);
    return bad;
}

/* Reweighting factor W_N (Luescher eq. 4.1): src/QED2dRHMC.jl:210-239.  The
 * j-th Gaussian field X (drawn with CUDA.randn at :218 in the reference) comes
 * from the Philox stream with purpose 0x60 and spin sub-counter 2j + c. */
double orc_reweighting_factor(const double *U, double am0, int n, const double *mu,
                              const double *rho, double A, double r_b, int N, int taylor,
                              int maxiter, double tol, uint64_t seed, uint32_t chain,
                              uint32_t traj, int L1, int L2, int bc, int *status)
{
    long V = (long)L1 * L2, m = 4 * V;
    double *X = (double *)malloc(sizeof(double) * m);
    double *ZpX = (double *)malloc(sizeof(double) * m);
    double *tmp = (double *)malloc(sizeof(double) * m);
    double W = 0.0;
    *status = 0;
    for (int j = 0; j < N; ++j) {
        double argexp = 0.0, Tfactor = 0.5;
        orc_fill_cgauss(X, L1, L2, seed, chain, traj, 0x60u, 2u * (uint32_t)j);
        memcpy(tmp, X, sizeof(double) * m);
        for (int i = 1; i <= taylor; ++i) {
            *status |= orc_luscher_z(ZpX, U, tmp, am0, n, mu, rho, A, r_b, maxiter, tol, L1, L2, bc);
            double re, im;
            orc_dot(X, ZpX, 2 * V, &re, &im);
            argexp += Tfactor * re;
            /* (-1)*(2i+1)/2/factorial(i+1)*factorial(i), evaluated left to right */
            double fi = 1.0, fi1;
            for (int q = 2; q <= i; ++q) fi *= q;
            fi1 = fi * (i + 1);
            Tfactor *= (double)(-1 * (2 * i + 1)) / 2.0 / fi1 * fi;
            memcpy(tmp, ZpX, sizeof(double) * m);
        }
        W += exp(argexp) / N;
    }
    free(X); free(ZpX); free(tmp);
    return W;
}

/* Lattice doubling: src/QED2dFields.jl:21-50 (reflect = conjugated copies). */
void orc_unfold(double *Uout_, const double *Uin_, int L1, int L2, int reflect)
{
    cplx *Uout = (cplx *)Uout_;
    const cplx *Uin = (const cplx *)Uin_;
    int M1 = 2 * L1, M2 = 2 * L2;
    for (int d = 0; d < 2; ++d)
        for (int i2 = 0; i2 < L2; ++i2)
            for (int i1 = 0; i1 < L1; ++i1) {
                cplx u = Uin[IDX(i1, i2, d, L1, L2)];
                cplx uc = reflect ? conj(u) : u;
                Uout[IDX(i1, i2, d, M1, M2)] = u;
                Uout[IDX(i1 + L1, i2, d, M1, M2)] = uc;
                Uout[IDX(i1, i2 + L2, d, M1, M2)] = uc;
                Uout[IDX(i1 + L1, i2 + L2, d, M1, M2)] = u;
            }
}

int orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
