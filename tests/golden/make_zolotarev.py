"""Writes tests/golden/zolotarev.json: Zolotarev coefficients of the reference's
get_rhmc_params (src/QED2dRHMC.jl:2-121) evaluated with mpmath at 60 digits.

Inputs are reproduced exactly as the reference forms them in FP64:
eps = (r_a / r_b)^2, k = sqrt(1 - eps), m = k^2 (each rounded to double), and
the elliptic functions are then evaluated at that double m with mpmath's
ellipk / ellipfun (parameter convention, as Elliptic.jl).  Run:
    python tests/golden/make_zolotarev.py
"""
import json
import math
import os

import mpmath as mp

mp.mp.dps = 60

CASES = [
    # (n, r_a, r_b)
    (1, math.sqrt(0.1), math.sqrt(26.0)),     # test/rhmc_pf_generation.jl:25-27
    (5, math.sqrt(0.1), math.sqrt(26.0)),
    (2, math.sqrt(0.3), math.sqrt(7.5)),
    (10, math.sqrt(0.02), math.sqrt(12.0)),
    (16, math.sqrt(0.005), math.sqrt(17.0)),  # 16 poles on [0.005, 17], the RHMC spectrum scale
    (8, math.sqrt(1.0e-4), math.sqrt(8.0)),
]


def coeffs(n, r_a, r_b):
    eps = (r_a / r_b) * (r_a / r_b)
    k = math.sqrt(1.0 - eps)
    m = k * k
    M = mp.mpf(m)
    K = mp.ellipk(M)
    v = K / (2 * n + 1)
    sn = {r: mp.ellipfun("sn", r * v, m=M) for r in range(1, 2 * n + 2)}
    cn = {r: mp.ellipfun("cn", r * v, m=M) for r in range(1, 2 * n + 2)}
    a = {r: cn[r] ** 2 / sn[r] ** 2 for r in sn}
    c = {r: sn[r] ** 2 for r in sn}
    rb = mp.mpf(r_b)
    mu = [rb * mp.sqrt(a[2 * j]) for j in range(1, n + 1)]
    nu = [rb * mp.sqrt(a[2 * j - 1]) for j in range(1, n + 1)]
    rho = []
    for j in range(n):
        res = nu[j] ** 2 - mu[j] ** 2
        for l in range(n):
            if l != j:
                res *= (nu[l] ** 2 - mu[j] ** 2) / (mu[l] ** 2 - mu[j] ** 2)
        rho.append(res)
    d = mp.mpf(k) ** (2 * n + 1)
    for i in range(1, 2 * n + 2, 2):
        d *= c[i] ** 2
    delta = d ** 2 / (1 + mp.sqrt(1 - d ** 2)) ** 2
    A = 2 / (1 + mp.sqrt(1 - d ** 2))
    for i in range(1, 2 * n, 2):
        A *= c[i]
    for i in range(2, 2 * n + 1, 2):
        A /= c[i]
    f = lambda x: float(x)
    return {"n": n, "r_a": r_a, "r_b": r_b, "eps": eps, "m": m, "K": f(K),
            "sn": [f(sn[r]) for r in range(1, 2 * n + 2)],
            "cn": [f(cn[r]) for r in range(1, 2 * n + 2)],
            "mu": [f(x) for x in mu], "nu": [f(x) for x in nu], "rho": [f(x) for x in rho],
            "delta": f(delta), "A": f(A)}


def main():
    out = [coeffs(*c) for c in CASES]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "zolotarev.json")
    with open(path, "w") as fh:
        json.dump({"generator": "tests/golden/make_zolotarev.py (mpmath %s, 60 digits)"
                   % mp.__version__, "cases": out}, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
