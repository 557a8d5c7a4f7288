// qed2d_fields.h -- per-site gauge-field and force building blocks shared by
// the single-lattice kernels (qed2d_kernels.hip) and the batched ensembles
// (qed2d_batch.hip).
#pragma once

#include "qed2d_march.h"

namespace qed2d {

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

// Fermion force z_mu at site x from the values it needs: the links U1, U2 of
// x (U2 with the fermion boundary sign already applied), X and Y at x, at
// x + 1 (a) and at x + 2 (b).  Returns Re(i z1), Re(i z2)
// (src/QED2dDirac.jl:172-199).
__device__ __forceinline__ void ferm_force_v(double2 U1, double2 U2, double2 X1, double2 X2,
                                             double2 Y1, double2 Y2, double2 Xa1, double2 Xa2,
                                             double2 Ya1, double2 Ya2, double2 Xb1, double2 Xb2,
                                             double2 Yb1, double2 Yb2, double &f1, double &f2)
{
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

// The same at site (i1, i2), loading what it needs.
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
    double2 U2 = U[V + s];
    if ((L.bc & 1) && i2 == L.L2 - 1) U2 = cneg<double>(U2);
    ferm_force_v(U[s], U2, X[s], X[V + s], Y[s], Y[V + s], X[s1], X[V + s1], Y[s1], Y[V + s1],
                 X[s2], X[V + s2], Y[s2], Y[V + s2], f1, f2);
}

// Gather form of the gauge force + fermion force + momentum update, one pass:
//   F(x,1) = -v(x) + v(x-2),  F(x,2) = v(x) - v(x-1),  v = beta Im P
// (the sums that src/QED2dAction.jl:47-50 scatters into frc1/frc2), then
//   mom += eps ((F_gauge) + cf F_fermion + facc)     (src/QED2dHMC.jl:478,554)
__device__ __forceinline__ void force_mom_site(const Latt &L, const double2 *__restrict__ U,
                                               const double2 *__restrict__ X,
                                               const double2 *__restrict__ Y,
                                               const double *__restrict__ facc,
                                               double *__restrict__ mom, double beta, double eps,
                                               double cf, long s)
{
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

}  // namespace qed2d
