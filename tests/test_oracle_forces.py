"""Pins the oracle's forces (restating src/QED2dDirac.jl:172-199 and
src/QED2dAction.jl:34-53) by central finite differences of the actions,
computed with dense linear algebra independently of the oracle's solver and
stencil.  These restate the reference's own FD force checks
(tests/gauge_forcetest.jl:35-70, tests/dev-fermion_forcetest.jl:51-114,
test/rhmc_pf_forcetest.jl:58-93), with the sign convention F = -dS/dtheta."""
import numpy as np
import pytest

from dense import q_dense, random_links, random_spinor, vec


def plaq_np(U):
    """Plaquettes by np.roll, independent of the oracle."""
    U1, U2 = U[0], U[1]
    return U1 * np.roll(U2, -1, axis=1) * np.conj(np.roll(U1, -1, axis=0) * U2)


def gauge_action_np(U, beta):
    return beta * np.sum(1.0 - plaq_np(U).real)


def twist(U, mu, i1, i2, d):
    V = U.copy()
    V[mu, i2, i1] *= np.exp(1j * d)
    return V


@pytest.mark.parametrize("link", [(0, 1, 3), (1, 4, 0), (0, 5, 5), (1, 0, 5)])
def test_gauge_force_fd(O, link):
    mu, i1, i2 = link
    rng = np.random.default_rng(17)
    L1, L2, beta = 6, 6, 3.7
    U = random_links(L1, L2, rng, unitary=True)
    f1, f2 = O.krnl_force(U, beta)
    F = (f1 + f2)[mu, i2, i1]
    h = 1e-5
    dS = (gauge_action_np(twist(U, mu, i1, i2, h), beta) -
          gauge_action_np(twist(U, mu, i1, i2, -h), beta)) / (2 * h)
    assert abs(F + dS) < 1e-8 * max(1.0, abs(F))


def test_gauge_force_scatter_writes_each_element_once(O):
    """krnl_force! scatters into frc2 (src/QED2dAction.jl:47-50); every
    element of frc1 and frc2 must be written exactly once per call."""
    L1, L2 = 5, 4
    U = np.ones((2, L2, L1), complex) * np.exp(0.3j)
    f1 = np.full((2, L2, L1), np.nan)
    f2 = np.full((2, L2, L1), np.nan)
    O.lib.orc_krnl_force(O._ptr(f1), O._ptr(f2), O._ptr(U), 1.0, L1, L2)
    assert not np.isnan(f1).any() and not np.isnan(f2).any()


def _fermion_action(U, phi, am0, mu2=0.0, bc=0):
    Qd = q_dense(U, am0, bc)
    A = Qd @ Qd + mu2 * np.eye(Qd.shape[0])
    return np.vdot(vec(phi), np.linalg.solve(A, vec(phi))).real


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("link", [(0, 1, 2), (1, 3, 3), (1, 2, 0)])
def test_fermion_force_fd(O, bc, link):
    mu, i1, i2 = link
    rng = np.random.default_rng(5 + bc)
    L1, L2, am0 = 4, 4, 0.3
    U = random_links(L1, L2, rng)           # non-unitary, as the reference tests
    phi = random_spinor(L1, L2, rng)
    Qd = q_dense(U, am0, bc)
    X = np.linalg.solve(Qd @ Qd, vec(phi)).reshape(2, L2, L1)
    Y = O.gamm5Dw(U, X, am0, bc)
    F = O.tr_dQwdU(U, X, Y, bc)[mu, i2, i1]
    h = 1e-5
    dS = (_fermion_action(twist(U, mu, i1, i2, h), phi, am0, bc=bc) -
          _fermion_action(twist(U, mu, i1, i2, -h), phi, am0, bc=bc)) / (2 * h)
    assert abs(F + dS) < 1e-7 * max(1.0, abs(F))


def test_rhmc_force_fd(O):
    """test/rhmc_pf_forcetest.jl:58-93: sum_i rho_i tr_dQwdU(X_i, Q X_i),
    X_i = (Q^2 + mu_i^2)^-1 phi, against the FD of phi^dag MultiCG(phi)."""
    rng = np.random.default_rng(1234)
    L1, L2, am0 = 4, 4, 0.5
    U = random_links(L1, L2, rng)
    phi = random_spinor(L1, L2, rng)
    z = O.zolotarev(3, np.sqrt(0.2), np.sqrt(30.0))
    Qd = q_dense(U, am0)
    F = np.zeros((2, L2, L1))
    for m, r in zip(z["mu"], z["rho"]):
        X = np.linalg.solve(Qd @ Qd + m * m * np.eye(Qd.shape[0]), vec(phi)).reshape(2, L2, L1)
        F += r * O.tr_dQwdU(U, X, O.gamm5Dw(U, X, am0))

    def S(Uv):
        return sum(r * _fermion_action(Uv, phi, am0, m * m) for m, r in zip(z["mu"], z["rho"]))

    h = 1e-5
    for (mu, i1, i2) in [(0, 1, 2), (1, 2, 3)]:
        dS = (S(twist(U, mu, i1, i2, h)) - S(twist(U, mu, i1, i2, -h))) / (2 * h)
        assert abs(F[mu, i2, i1] + dS) < 1e-7 * max(1.0, abs(dS))


def test_fused_force_matches_parts(O):
    """mom + eps (frc1 + frc2 + frc) as src/QED2dHMC.jl:478, assembled by the
    oracle's update_momenta from its parts."""
    rng = np.random.default_rng(3)
    L1, L2, am0, beta = 6, 6, 0.4, 2.5
    U = random_links(L1, L2, rng, unitary=True)
    phi = random_spinor(L1, L2, rng)
    mom = rng.standard_normal((2, L2, L1))
    m1 = mom.copy()
    X = np.zeros_like(phi)
    g5 = np.zeros_like(phi)
    p = O.latt(L1, L2, beta)
    it = O.lib.orc_update_momenta_nf2(O._ptr(m1), O._ptr(U), O._ptr(X), O._ptr(phi), O._ptr(g5),
                                      am0, 0.1, 1000, 1e-26, O.C.byref(p))
    assert it > 0
    Qd = q_dense(U, am0)
    Xd = np.linalg.solve(Qd @ Qd, vec(phi)).reshape(2, L2, L1)
    f1, f2 = O.krnl_force(U, beta)
    ff = O.tr_dQwdU(U, Xd, O.gamm5Dw(U, Xd, am0))
    np.testing.assert_allclose(m1, mom + 0.1 * (f1 + f2 + ff), rtol=0, atol=1e-10)
