// qed2d_zolo.cpp -- Zolotarev optimal rational approximation of y^(-1/2) on
// [r_a^2, r_b^2]: the coefficients of get_rhmc_params
// (reference src/QED2dRHMC.jl:2-121).
//
// The reference evaluates the complete elliptic integral K(m) and the Jacobi
// functions sn(u|m), cn(u|m) with Elliptic.jl 1.0.1 (parameter m = k^2).
// Here K comes from the arithmetic-geometric mean, K(m) = pi / (2 AGM(1,
// sqrt(1-m))) (Abramowitz & Stegun 17.6.1), and sn, cn from the descending
// Landen / AGM scheme (A&S 16.4): run the AGM recording c_n / a_n, set
// phi_N = 2^N a_N u and recur phi_{n-1} = (phi_n + asin(c_n/a_n sin phi_n))/2.
#include <cmath>
#include <vector>

#include "../../include/qed2d.h"

namespace {

double ellipk(double m)
{
    double a = 1.0, b = std::sqrt(1.0 - m);
    for (int i = 0; i < 64 && std::fabs(a - b) > 1e-17 * a; ++i) {
        double an = 0.5 * (a + b);
        b = std::sqrt(a * b);
        a = an;
    }
    return M_PI / (2.0 * a);
}

void sncn(double u, double m, double *sn, double *cn)
{
    double a[80], c[80];
    double b = std::sqrt(1.0 - m);
    a[0] = 1.0;
    c[0] = std::sqrt(m);
    int N = 0;
    while (N < 70 && std::fabs(c[N]) > 1e-17 * a[N]) {
        a[N + 1] = 0.5 * (a[N] + b);
        c[N + 1] = 0.5 * (a[N] - b);
        b = std::sqrt(a[N] * b);
        ++N;
    }
    double phi = std::ldexp(a[N] * u, N);
    for (int k = N; k >= 1; --k) phi = 0.5 * (phi + std::asin(c[k] / a[k] * std::sin(phi)));
    *sn = std::sin(phi);
    *cn = std::cos(phi);
}

struct Zolo {
    int n;
    double k, m, v;
    // sn^2 and cn^2 / sn^2 at u = r v for r = 1 .. 2n+1
    std::vector<double> c2, a2;
    Zolo(int n_, double eps) : n(n_)
    {
        k = std::sqrt(1.0 - eps);            // k(eps)      (:2)
        m = k * k;
        v = ellipk(m) / (2.0 * n + 1.0);     // v(n, eps)   (:3)
        c2.resize(2 * n + 2);
        a2.resize(2 * n + 2);
        for (int r = 1; r <= 2 * n + 1; ++r) {
            double s, c;
            sncn(r * v, m, &s, &c);
            a2[r] = (c * c) / (s * s);       // a_r         (:4)
            c2[r] = s * s;                   // c_r         (:5)
        }
    }
    double mu(int j, double r_b) const { return r_b * std::pow(a2[2 * j], 0.5); }      // (:8)
    double nu(int j, double r_b) const { return r_b * std::pow(a2[2 * j - 1], 0.5); }  // (:9)
};

}  // namespace

extern "C" int qed2d_zolotarev(int n, double r_a, double r_b, double *eps_out, double *A_out,
                               double *delta_out, double *rho, double *mu, double *nu)
{
    if (n < 1 || n > 256 || !(r_a > 0.0) || !(r_b > r_a) || !rho || !mu || !nu)
        return QED2D_EINVAL;
    const double eps = (r_a / r_b) * (r_a / r_b);   // src/QED2dRHMC.jl:95
    Zolo z(n, eps);
    // d(n, eps)  (:60-68)
    double d = std::pow(z.k, 2 * n + 1);
    for (int i = 1; i <= 2 * n + 1; i += 2) d *= z.c2[i] * z.c2[i];
    const double sq = std::sqrt(1.0 - d * d);
    const double delta = d * d / ((1.0 + sq) * (1.0 + sq));  // (:70)
    double A = 2.0 / (1.0 + sq);                               // (:72-84)
    for (int i = 1; i <= 2 * n - 1; i += 2) A *= z.c2[i];
    for (int i = 2; i <= 2 * n; i += 2) A *= 1.0 / z.c2[i];
    for (int j = 1; j <= n; ++j) {
        const double muj = z.mu(j, r_b), nuj = z.nu(j, r_b);
        mu[j - 1] = muj;
        nu[j - 1] = nuj;
        // rho_mu(j, 1, n, n, eps, r_b), Luescher eq. (3.5)  (:11-25)
        double res = nuj * nuj - muj * muj;
        for (int l = 1; l <= n; ++l) {
            if (l == j) continue;
            const double nul = z.nu(l, r_b), mul = z.mu(l, r_b);
            res *= (nul * nul - muj * muj) / (mul * mul - muj * muj);
        }
        rho[j - 1] = res;
    }
    if (eps_out) *eps_out = eps;
    if (A_out) *A_out = A;
    if (delta_out) *delta_out = delta;
    return QED2D_OK;
}
