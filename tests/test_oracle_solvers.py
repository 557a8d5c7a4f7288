"""Pins the oracle's solvers (src/QED2dSolver.jl, src/QED2dRHMC.jl:135-170,
261-319) against dense linear algebra on small lattices."""
import numpy as np
import pytest

from dense import q_dense, random_links, random_spinor, vec


def _setup(L1=6, L2=4, seed=0, am0=0.25):
    rng = np.random.default_rng(seed)
    U = random_links(L1, L2, rng)
    b = random_spinor(L1, L2, rng)
    Qd = q_dense(U, am0)
    return U, b, Qd, am0


@pytest.mark.parametrize("mu", [0.0, 0.7])
def test_cg_solves_dense_system(O, mu):
    U, b, Qd, am0 = _setup()
    A = Qd @ Qd + mu * mu * np.eye(Qd.shape[0])
    xd = np.linalg.solve(A, vec(b))
    x, it, st = O.cg(U, b, am0, 2000, 1e-26, mu=mu, op=1 if mu else 0)
    assert st == 0 and it > 0
    assert np.max(np.abs(vec(x) - xd)) < 1e-10 * np.max(np.abs(xd))


def test_cg_stopping_rule_is_squared_relative(O):
    """src/QED2dSolver.jl:16,29: stop when |r|^2 < eps |b|^2; the iteration
    before the returned one did not satisfy it; maxiter one short fails."""
    U, b, Qd, am0 = _setup(seed=3)
    eps = 1e-12
    x, it, st = O.cg(U, b, am0, 2000, eps)
    r = vec(b) - Qd @ Qd @ vec(x)
    assert st == 0
    assert np.vdot(r, r).real < 1e-11 * np.vdot(vec(b), vec(b)).real   # rounding slack
    x2, it2, st2 = O.cg(U, b, am0, it - 1, eps)
    assert st2 == 1


def test_cg_v2_guess(O):
    U, b, Qd, am0 = _setup(seed=4)
    x, it, st = O.cg(U, b, am0, 2000, 1e-24, mu=0.5, op=1)
    x2, it2, st2 = O.cg(U, b, am0, 2000, 1e-24, mu=0.5, op=1, x0=x)
    assert it2 == 0 and st2 == 0                     # src/QED2dSolver.jl:65-72
    x3, it3, st3 = O.cg(U, b, am0, 2000, 1e-24, mu=0.9, op=1, x0=x)
    A = Qd @ Qd + 0.81 * np.eye(Qd.shape[0])
    assert it3 > 0 and np.max(np.abs(A @ vec(x3) - vec(b))) < 1e-9


def test_multicg_and_R(O):
    U, b, Qd, am0 = _setup(seed=5)
    z = O.zolotarev(4, np.sqrt(0.1), np.sqrt(26.0))
    n = Qd.shape[0]
    M = np.eye(n, dtype=complex)
    for m, r in zip(z["mu"], z["rho"]):
        M = M + r * np.linalg.inv(Qd @ Qd + m * m * np.eye(n))
    so, bad = O.multicg(U, b, am0, z["mu"], z["rho"], 5000, 1e-26)
    assert bad == 0
    ref = M @ vec(b)
    assert np.max(np.abs(vec(so) - ref)) < 1e-10 * np.max(np.abs(ref))
    so, bad = O.R(U, b, am0, z, 5000, 1e-26)
    ref = (z["A"] / z["r_b"]) * ref
    assert np.max(np.abs(vec(so) - ref)) < 1e-10 * np.max(np.abs(ref))


def test_generate_pseudofermion_identity(O):
    """test/rhmc_pf_generation.jl:56-57: phi = prod (Q - i nu)(Q^2+nu^2)^-1 (Q + i mu) X
    gives phi^dag (1 + sum rho (Q^2 + mu^2)^-1) phi = X^dag X."""
    L1, L2 = 10, 10
    rng = np.random.default_rng(1234)
    U = random_links(L1, L2, rng)
    for am0 in (0.02, 1.2):
        z = O.zolotarev(1, np.sqrt(0.1), np.sqrt(26.0))
        X = random_spinor(L1, L2, rng)
        F, bad = O.generate_pseudofermion(U, X, am0, z, 10000, 1e-25)
        xi, bad2 = O.multicg(U, F, am0, z["mu"], z["rho"], 10000, 1e-25)
        assert bad == 0 and bad2 == 0
        assert abs(O.dot(xi, F).real - O.norm2(X)) < 1e-10 * O.norm2(X)


def test_generate_pseudofermion_dense(O):
    U, X, Qd, am0 = _setup(seed=8)
    z = O.zolotarev(3, np.sqrt(0.1), np.sqrt(26.0))
    n = Qd.shape[0]
    v = vec(X)
    for m, nu in zip(z["mu"], z["nu"]):
        v = (Qd + 1j * m * np.eye(n)) @ v
        v = np.linalg.solve(Qd @ Qd + nu * nu * np.eye(n), v)
        v = (Qd - 1j * nu * np.eye(n)) @ v
    F, bad = O.generate_pseudofermion(U, X, am0, z, 5000, 1e-26)
    assert np.max(np.abs(vec(F) - v)) < 1e-10 * np.max(np.abs(v))


def test_power_method_matches_dense_spectrum(O):
    rng = np.random.default_rng(21)
    L1, L2, am0 = 4, 4, 0.8
    U = random_links(L1, L2, rng)
    b0 = random_spinor(L1, L2, rng)
    Qd = q_dense(U, am0)
    ev = np.linalg.eigvalsh(Qd @ Qd)
    lmin, lmax = O.power_method(U, b0, am0, 4000)
    assert abs(lmax - ev[-1]) < 1e-8 * ev[-1]
    assert abs(lmin - ev[0]) < 1e-3 * ev[-1]
