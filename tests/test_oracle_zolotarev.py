"""Pins the Zolotarev coefficients (src/QED2dRHMC.jl:2-121; elliptic functions
of Elliptic.jl 1.0.1 in the reference, own AGM / Landen routines here) against
mpmath golden vectors (tests/golden/zolotarev.json, made by
tests/golden/make_zolotarev.py) and against the defining property of the
approximation: 1 - sqrt(y) R(y) equioscillates with amplitude delta on
[r_a^2, r_b^2].  Also the reference's operator-level bound
(test/rhmc_zolovarev_bound.jl:30-42)."""
import json
import os

import numpy as np
import pytest

from dense import q_dense, random_links, random_spinor, vec

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "zolotarev.json")))
CASES = GOLD["cases"]


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c['n']}")
def test_elliptic_functions(O, case):
    m = case["m"]
    assert rel(O.ellipk(m), case["K"]) < 1e-14
    v = case["K"] / (2 * case["n"] + 1)
    for r in range(1, 2 * case["n"] + 2):
        s, c = O.ellip_sncn(r * v, m)
        assert abs(s - case["sn"][r - 1]) < 1e-14
        # cn -> 0 at u = K; compare absolutely there
        assert abs(c - case["cn"][r - 1]) < 1e-14 * max(1.0, 1e3 * abs(case["cn"][r - 1]))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c['n']}")
def test_coefficients_vs_mpmath(O, case):
    z = O.zolotarev(case["n"], case["r_a"], case["r_b"])
    assert z["eps"] == case["eps"]
    assert rel(z["mu"], case["mu"]) < 1e-11
    assert rel(z["nu"], case["nu"]) < 1e-11
    assert rel(z["rho"], case["rho"]) < 1e-10
    assert rel(z["A"], case["A"]) < 1e-11
    assert rel(z["delta"], case["delta"]) < 1e-8


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c['n']}")
def test_equioscillation(O, case):
    """max over [r_a^2, r_b^2] of |1 - sqrt(y) R(y)| equals delta, with
    R(y) = (A / r_b) (1 + sum_j rho_j / (y + mu_j^2))."""
    z = O.zolotarev(case["n"], case["r_a"], case["r_b"])
    y = np.exp(np.linspace(np.log(case["r_a"] ** 2), np.log(case["r_b"] ** 2), 200001))
    R = np.ones_like(y)
    for m, r in zip(z["mu"], z["rho"]):
        R = R + r / (y + m * m)
    R *= z["A"] / z["r_b"]
    err = np.abs(1.0 - np.sqrt(y) * R)
    if z["delta"] > 1e-12:          # resolvable in FP64
        assert err.max() <= z["delta"] * (1 + 1e-4) + 1e-15
        assert err.max() >= z["delta"] * 0.99
        # 2n + 2 alternation points: count sign changes of the error
        e = 1.0 - np.sqrt(y) * R
        assert np.sum(np.diff(np.sign(e)) != 0) >= 2 * case["n"]
    else:
        assert err.max() < 1e-12


def test_operator_bound(O):
    """test/rhmc_zolovarev_bound.jl:30-42: |X - Q^2 R^2 X| / |X| < delta (2 + delta)
    with spectral bounds widened by 0.99 / 1.01 (here from the dense spectrum)."""
    rng = np.random.default_rng(20)
    L1, L2, am0 = 6, 6, 0.2
    U = random_links(L1, L2, rng)
    Qd = q_dense(U, am0)
    ev = np.linalg.eigvalsh(Qd @ Qd)
    z = O.zolotarev(5, np.sqrt(ev[0] * 0.99), np.sqrt(ev[-1] * 1.01))
    X = random_spinor(L1, L2, rng)
    F, _ = O.R(U, X, am0, z, 10000, 1e-25)
    F2, _ = O.R(U, F, am0, z, 10000, 1e-25)
    Xf = O.gamm5Dw_sqr(U, F2, am0)
    dev = np.sqrt(O.norm2(X - Xf) / O.norm2(X))
    assert dev < z["delta"] * (2 + z["delta"])
