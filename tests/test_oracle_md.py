"""Pins the oracle's link update, observables, integrators and HMC drivers
(src/QED2dAction.jl, src/QED2dHMC.jl) by closed forms and invariants:
cold start, constant-flux fields with integer charge, reversibility
(test/rhmc_hmc_reversibility.jl:41-50 restated for all three HMC flavours),
the energy-violation scaling dH ~ eps^2 (leapfrog) and eps^4 (OMF4)."""
import numpy as np
import pytest

from dense import q_dense, random_links, vec


def test_updt_fld_closed_form(O):
    rng = np.random.default_rng(0)
    U = random_links(8, 6, rng)
    mom = rng.standard_normal((2, 6, 8))
    np.testing.assert_allclose(O.updt_fld(U, mom, 0.37), np.exp(0.37j * mom) * U, rtol=1e-15,
                               atol=1e-15)


def test_cold_start(O):
    U = np.ones((2, 12, 10), complex)
    assert O.plaquette(U) == 1.0 and O.action(U, 4.0) == 0.0 and O.qtop(U) == 0.0
    f1, f2 = O.krnl_force(U, 4.0)
    assert np.all(f1 + f2 == 0.0)


def flux_links(L1, L2, q):
    """Every plaquette angle = 2 pi q / V (mod 2 pi): U1 = 1 except on the
    last column, U2(i1, i2) = exp(2 pi i q i1 / V)."""
    V = L1 * L2
    th = np.zeros((2, L2, L1))
    i1 = np.arange(L1)[None, :]
    i2 = np.arange(L2)[:, None]
    th[1] = 2 * np.pi * q * i1 / V + 0 * i2
    th[0][:, L1 - 1] = -2 * np.pi * q * np.arange(L2) * L1 / V
    return np.exp(1j * th)


@pytest.mark.parametrize("q", [-3, -1, 0, 1, 2, 5])
def test_constant_flux_charge(O, q):
    L1, L2 = 16, 12
    U = flux_links(L1, L2, q)
    qd = O.qtop_field(U)
    np.testing.assert_allclose(qd, 2 * np.pi * q / (L1 * L2), atol=1e-12)
    assert round(O.qtop(U)) == q and abs(O.qtop(U) - q) < 1e-10
    assert abs(O.plaquette(U) - np.cos(2 * np.pi * q / (L1 * L2))) < 1e-14


def _pure_dH(O, eps, ns, integ):
    L = 16
    U = O.hot_start(L, L, 3)
    # a smoother start than a hot one: a few OMF4 trajectories
    p = O.latt(L, L, 2.0)
    for t in range(3):
        O.hmc_pure(U, 0.1, 10, p, 3, t)
    mom = O.momenta(L, L, 3, 0, 100)
    h0 = O.hamiltonian(mom, U, 2.0)
    if integ == "omf4":
        O.omf4_pure(mom, U, eps, ns, p)
    else:
        O.leapfrog_pure(mom, U, eps, ns, p)
    return abs(O.hamiltonian(mom, U, 2.0) - h0)


def test_energy_scaling_leapfrog(O):
    r = _pure_dH(O, 0.1, 10, "lf") / _pure_dH(O, 0.05, 20, "lf")
    assert 3.0 < r < 5.0


def test_energy_scaling_omf4(O):
    r = _pure_dH(O, 0.2, 5, "omf4") / _pure_dH(O, 0.1, 10, "omf4")
    assert 10.0 < r < 22.0


def test_pure_gauge_reversibility(O):
    L = 12
    U = O.hot_start(L, L, 9)
    U0 = U.copy()
    mom = O.momenta(L, L, 9, 0, 0)
    p = O.latt(L, L, 3.0)
    O.omf4_pure(mom, U, 0.1, 10, p)
    mom *= -1
    O.omf4_pure(mom, U, 0.1, 10, p)
    assert np.max(np.abs(U - U0)) < 1e-12


@pytest.mark.parametrize("integ", ["lf", "omf4"])
def test_nf2_reversibility(O, integ):
    L = 8
    U = O.hot_start(L, L, 4)
    U0 = U.copy()
    am0 = 0.5
    mom = O.momenta(L, L, 4, 0, 0)
    X = O.pseudofermion(L, L, 4, 0, 0)
    F = O.gamm5Dw(U, X, am0)
    S0 = O.action(U, 4.0)
    p = O.latt(L, L, 4.0)
    Xs = np.zeros_like(X)
    run = O.omf4_nf2 if integ == "omf4" else O.leapfrog_nf2
    assert run(mom, U, Xs, F, am0, 0.05, 6, 10000, 1e-28, p) > 0
    mom *= -1
    assert run(mom, U, Xs, F, am0, 0.05, 6, 10000, 1e-28, p) > 0
    assert np.max(np.abs(U - U0)) < 1e-11
    assert abs(O.action(U, 4.0) - S0) < 1e-10


def test_nf2_energy_scaling(O):
    L = 8
    am0 = 0.5
    p = O.latt(L, L, 4.0)

    def dH(eps, ns):
        U = O.hot_start(L, L, 6)
        mom = O.momenta(L, L, 6, 0, 0)
        X = O.pseudofermion(L, L, 6, 0, 0)
        F = O.gamm5Dw(U, X, am0)
        h0 = O.hamiltonian(mom, U, 4.0) + O.norm2(X)
        Xs = np.zeros_like(X)
        O.leapfrog_nf2(mom, U, Xs, F, am0, eps, ns, 10000, 1e-28, p)
        return abs(O.hamiltonian(mom, U, 4.0) + O.dot(Xs, F).real - h0)

    r = dH(0.1, 5) / dH(0.05, 10)
    assert 3.0 < r < 5.0


def test_nf2_pseudofermion_action_at_start(O):
    """hini's X^dag X equals phi^dag (Q^2)^-1 phi for phi = Q X
    (src/QED2dHMC.jl:70-82): zero-step MD has dH = 0 up to the CG residual."""
    L = 6
    U = O.hot_start(L, L, 2)
    X = O.pseudofermion(L, L, 2, 0, 0)
    F = O.gamm5Dw(U, X, 0.3)
    x, it, st = O.cg(U, F, 0.3, 5000, 1e-28)
    assert abs(O.dot(x, F).real - O.norm2(X)) < 1e-10 * O.norm2(X)


def test_hmc_nf2_driver_accounting(O):
    """hmc_nf2 reproduces hini = H + |X|^2 with the Philox random streams and
    accepts / rejects on the Metropolis uniform of the same trajectory."""
    L = 8
    p = O.latt(L, L, 4.0)
    U = O.hot_start(L, L, 1234)
    U0 = U.copy()
    hi, hf, acc, it, rc = O.hmc_nf2(U, 0.3, 0.1, 5, 10000, 1e-24, p, 1234, 7)
    mom = O.momenta(L, L, 1234, 0, 7)
    X = O.pseudofermion(L, L, 1234, 0, 7)
    assert rc == 0
    assert abs(hi - (O.hamiltonian(mom, U0, 4.0) + O.norm2(X))) < 1e-10 * abs(hi)
    u = O.metropolis_uniform(1234, 0, 7)
    pacc = np.exp(-(hf - hi))
    assert acc == (1 if pacc >= 1.0 or pacc > u else 0)
    if not acc:
        assert np.array_equal(U, U0)


@pytest.mark.parametrize("L,integ", [(8, 0), (24, 0), (24, 1), (64, 0)])
def test_hmc_nf2_extended_precision_dh(O, L, integ):
    """last_dh (energies in extended precision) equals hfin - hini where the
    doubles resolve it: H is ~1e2 .. 1e4 here, so hfin - hini carries at most
    ~1e-12 of rounding, while a dropped or mis-signed term of H (kinetic,
    action, X^dag X / Re X^dag F) would differ by O(1)."""
    p = O.latt(L, L, 4.0)
    U = O.hot_start(L, L, 99 + L)
    for t in range(3):
        hi, hf, acc, it, rc = O.hmc_nf2(U, 0.3, 0.1, 4, 10000, 1e-24, p, 4321, t,
                                        integrator=integ)
        assert rc == 0
        dh = O.last_dh()
        assert abs(dh - (hf - hi)) < 4e-16 * abs(hi) + 1e-13, (t, dh, hf - hi)
        assert dh != 0.0


def test_rhmc_reversibility(O):
    """test/rhmc_hmc_reversibility.jl: 10x10, am0 = 10, n = 5 poles on the
    power-method bounds, 100 steps of eps = 0.01, tol 1e-25: |dS_gauge| < 1e-10."""
    L = 10
    rng = np.random.default_rng(1234)
    U = random_links(L, L, rng)
    b0 = (rng.standard_normal((2, L, L)) + 1j * rng.standard_normal((2, L, L))) / np.sqrt(2)
    lmin, lmax = O.power_method(U, b0, 10.0, 1000)
    z = O.zolotarev(5, np.sqrt(lmin), np.sqrt(lmax))
    p = O.latt(L, L, 6.05)
    S0 = O.action(U, 6.05)
    hi, hf, acc, rc = O.hmc_rhmc(U, [10.0], [z], 0.01, 100, 10000, 1e-25, p, 1234, 0,
                                 reversibility_test=1)
    assert rc == 0 and acc == -1
    assert abs(O.action(U, 6.05) - S0) < 1e-10


def test_unfold(O):
    rng = np.random.default_rng(0)
    U = random_links(5, 3, rng)
    W = O.unfold(U, False)
    for a in (0, 1):
        for b in (0, 1):
            assert np.array_equal(W[:, b * 3:(b + 1) * 3, a * 5:(a + 1) * 5], U)
    R = O.unfold(U, True)
    assert np.array_equal(R[:, 0:3, 5:10], np.conj(U))
    assert np.array_equal(R[:, 3:6, 0:5], np.conj(U))
    assert np.array_equal(R[:, 3:6, 5:10], U)
    # the doubled periodic field has the same plaquette
    assert abs(O.plaquette(W) - O.plaquette(U)) < 1e-14
