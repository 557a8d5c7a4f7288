"""Pins the oracle's reweighting restatement (src/QED2dRHMC.jl:174-239) on 6x6
lattices where Q^2 is a dense 72x72 matrix:

* Z = Q^2 R(Q^2)^2 - 1 from the eigen-decomposition of the dense Q^2 and
  R(y) = (A / r_b)(1 + sum_j rho_j / (y + mu_j^2));
* the Taylor factors are the binomial coefficients -C(-1/2, i) (scipy), so
  the truncated series sum_i T_i X^dag Z^i X converges to
  X^dag (1 - (1 + Z)^{-1/2}) X, computed in closed form;
* the reference's own bounds (test/rhmc_zolovarev_bound.jl:59-85):
  |(X, Z^p X)| < 2V (2 delta)^p and exp(-2V delta) < W_1 < exp(2V delta)."""
import numpy as np
import pytest
from scipy.special import binom

from dense import q_dense, random_links, vec

L1 = L2 = 6
AM0 = 0.2


def setup(seed, n, widen=(0.99, 1.01)):
    rng = np.random.default_rng(seed)
    U = random_links(L1, L2, rng, unitary=True)
    Qd = q_dense(U, AM0)
    Q2 = Qd @ Qd
    Q2 = 0.5 * (Q2 + Q2.conj().T)
    ev, V = np.linalg.eigh(Q2)
    return U, ev, V


def zdense(ev, V, z):
    R = z["A"] / z["r_b"] * (1 + sum(r / (ev + m * m) for m, r in zip(z["mu"], z["rho"])))
    zz = ev * R * R - 1.0
    return zz, V


def apply(zz, V, x, p=1):
    return V @ (zz ** p * (V.conj().T @ x))


@pytest.mark.parametrize("n,p", [(2, 1), (2, 3), (4, 2)])
def test_luscher_z_dense(O, n, p):
    U, ev, V = setup(1, n)
    z = O.zolotarev(n, np.sqrt(ev[0] * 0.99), np.sqrt(ev[-1] * 1.01))
    zz, V = zdense(ev, V, z)
    rng = np.random.default_rng(5)
    X = (rng.standard_normal((2, L2, L1)) + 1j * rng.standard_normal((2, L2, L1)))
    Z, bad = O.luscher_z(U, X, AM0, z, 10000, 1e-26, p=p)
    assert bad == 0
    ref = apply(zz, V, vec(X), p)
    # |Z^p X| is O(delta^p |X|) and Q^2 R^2 X - X cancels to that size, so the
    # error is measured against the field itself; relative to Z^p X loosely
    assert np.linalg.norm(vec(Z) - ref) / np.linalg.norm(vec(X)) < 1e-12
    assert np.linalg.norm(vec(Z) - ref) / np.linalg.norm(ref) < 1e-5


def test_luscher_z_trace_bound(O):
    """test/rhmc_zolovarev_bound.jl:59-67 with 3 poles (at 10 poles delta is
    1e-14 on a 6x6 spectrum and Z^2 X sits at the rounding floor)."""
    U, ev, V = setup(2, 3)
    z = O.zolotarev(3, np.sqrt(ev[0] * 0.99), np.sqrt(ev[-1] * 1.01))
    X = O.cgauss(L1, L2, 77, 0, 0, 0x60)
    for p in range(1, 4):
        Z, _ = O.luscher_z(U, X, AM0, z, 100000, 1e-25, p=p)
        tr = np.vdot(vec(X), vec(Z)).real
        assert abs(tr) < 2 * L1 * L2 * (2 * z["delta"]) ** p


@pytest.mark.parametrize("N,taylor,n", [(1, 1, 2), (3, 4, 2), (2, 5, 3)])
def test_reweighting_factor_dense(O, N, taylor, n):
    U, ev, V = setup(3, n)
    z = O.zolotarev(n, np.sqrt(ev[0] * 0.99), np.sqrt(ev[-1] * 1.01))
    zz, V = zdense(ev, V, z)
    seed, chain, traj = 4242, 1, 7
    W, st = O.reweighting_factor(U, AM0, z, N, taylor, 10000, 1e-26, seed, chain, traj)
    assert st == 0
    ref = 0.0
    for j in range(N):
        x = vec(O.cgauss(L1, L2, seed, chain, traj, 0x60, 2 * j))
        a = sum(-binom(-0.5, i) * np.vdot(x, apply(zz, V, x, i)).real
                for i in range(1, taylor + 1))
        ref += np.exp(a) / N
    assert abs(W - ref) < 1e-11 * abs(ref)
    assert abs(W - 1.0) > 1e-6          # the correction is visible at n = 2, 3


def test_reweighting_series_limit(O):
    """At high Taylor order the series reaches exp(X^dag (1 - (1+Z)^{-1/2}) X)."""
    U, ev, V = setup(4, 2)
    z = O.zolotarev(2, np.sqrt(ev[0] * 0.99), np.sqrt(ev[-1] * 1.01))
    zz, V = zdense(ev, V, z)
    assert np.max(np.abs(zz)) < 0.2
    W, _ = O.reweighting_factor(U, AM0, z, 1, 14, 10000, 1e-26, 9, 0, 0)
    x = vec(O.cgauss(L1, L2, 9, 0, 0, 0x60, 0))
    ref = np.exp(np.vdot(x, apply(1.0 - (1.0 + zz) ** -0.5, V, x)).real)
    assert abs(W - ref) < 1e-10 * ref


@pytest.mark.parametrize("n", [1, 4, 7])
def test_reweighting_bound(O, n):
    """test/rhmc_zolovarev_bound.jl:76-85."""
    U, ev, V = setup(5, n)
    z = O.zolotarev(n, np.sqrt(ev[0] * 0.99), np.sqrt(ev[-1] * 1.01))
    W, _ = O.reweighting_factor(U, AM0, z, 1, 5, 100000, 1e-25, 1234, 0, n)
    V_ = L1 * L2
    assert np.exp(-2 * V_ * z["delta"]) < W < np.exp(2 * V_ * z["delta"])
