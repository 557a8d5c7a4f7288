"""Pins the oracle's Wilson operator (oracle/qed2d_oracle.c, restating
src/QED2dDirac.jl) against the dense gamma-matrix Dirac matrix, the free-field
plane-wave spectrum (tests/Dw-eigenvalues.jl:21-35, src/QED2dUtils.jl:2-11),
gamma5-hermiticity and gauge covariance."""
import numpy as np
import pytest

from dense import (field, gamma5_dense, dirac_dense, q_dense, random_links, random_spinor,
                   vec)

SIZES = [(4, 4), (6, 4), (5, 3)]


@pytest.mark.parametrize("L1,L2", SIZES)
@pytest.mark.parametrize("bc", [0, 1])
def test_q_matches_dense(O, L1, L2, bc):
    rng = np.random.default_rng(11 + L1 * 7 + L2 + bc)
    U = random_links(L1, L2, rng)
    si = random_spinor(L1, L2, rng)
    am0 = 0.3
    so = O.gamm5Dw(U, si, am0, bc)
    ref = field(q_dense(U, am0, bc) @ vec(si), L1, L2)
    assert np.max(np.abs(so - ref)) < 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("L1,L2", SIZES[:2])
def test_dw_gamma5_sqr_forms(O, L1, L2):
    rng = np.random.default_rng(5)
    U = random_links(L1, L2, rng)
    si = random_spinor(L1, L2, rng)
    am0, mu = 0.2, 0.7
    Qd = q_dense(U, am0)
    Dd = dirac_dense(U, am0)
    v = vec(si)
    scale = np.max(np.abs(Qd @ Qd @ Qd @ Qd @ v))
    np.testing.assert_allclose(vec(O.Dw(U, si, am0)), Dd @ v, atol=1e-13 * np.max(np.abs(Dd @ v)))
    np.testing.assert_allclose(vec(O.gamm5(si)), gamma5_dense(L1 * L2) @ v, atol=0)
    np.testing.assert_allclose(vec(O.gamm5Dw_sqr(U, si, am0)), Qd @ Qd @ v, atol=1e-13 * scale)
    np.testing.assert_allclose(vec(O.gamm5Dw_sqr_musq(U, si, am0, mu)),
                               (Qd @ Qd + mu ** 2 * np.eye(len(v))) @ v, atol=1e-13 * scale)
    np.testing.assert_allclose(vec(O.gamm5Dw_sqr_sqr(U, si, am0)), Qd @ Qd @ Qd @ Qd @ v,
                               atol=1e-13 * scale)
    np.testing.assert_allclose(vec(O.gamm5Dw_sqr_sqr_musq(U, si, am0, mu)),
                               Qd @ Qd @ Qd @ Qd @ v + mu ** 2 * v, atol=1e-13 * scale)


def test_q_hermitian_for_nonunitary_links(O):
    """Q = gamma5 Dw is Hermitian for any complex U (backward hop uses conj)."""
    rng = np.random.default_rng(3)
    L1, L2 = 8, 6
    U = random_links(L1, L2, rng)
    a, b = random_spinor(L1, L2, rng), random_spinor(L1, L2, rng)
    lhs = np.vdot(vec(a), vec(O.gamm5Dw(U, b, 0.1)))
    rhs = np.vdot(vec(O.gamm5Dw(U, a, 0.1)), vec(b))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("n1,n2", [(5, 5), (1, 3), (0, 2)])
def test_free_field_plane_wave(O, n1, n2):
    """tests/Dw-eigenvalues.jl:21-35: with U = 1 the plane wave of
    src/QED2dUtils.jl:2-11 is an eigenvector of Dw with eigenvalue
    m + 2 sum sin^2(k/2) +- i sqrt(sum sin^2 k)."""
    L = 20
    am0 = 10.0
    k1, k2 = 2 * np.pi * n1 / L, 2 * np.pi * n2 / L
    U = np.ones((2, L, L), complex)
    i1 = np.arange(L)[None, :]
    i2 = np.arange(L)[:, None]
    ph = np.exp(1j * (k1 * i1 + k2 * i2))
    s = np.sin(k1) ** 2 + np.sin(k2) ** 2
    b = np.zeros((2, L, L), complex)
    if s > 0:
        b[0] = 1j * np.sqrt(complex(-s)) / (np.sin(k1) + 1j * np.sin(k2)) * ph
    else:
        b[0] = ph
    b[1] = ph
    bn = O.Dw(U, b, am0)
    lam = np.vdot(vec(b), vec(bn)) / np.vdot(vec(b), vec(b))
    lam1 = am0 + 2 * (np.sin(k1 / 2) ** 2 + np.sin(k2 / 2) ** 2) + 1j * np.sqrt(s)
    lam2 = am0 + 2 * (np.sin(k1 / 2) ** 2 + np.sin(k2 / 2) ** 2) - 1j * np.sqrt(s)
    assert min(abs(lam - lam1), abs(lam - lam2)) < 1e-12 * abs(lam1)
    assert np.max(np.abs(bn / lam - b)) < 1e-12


def test_gauge_covariance(O):
    """Q[U^g] (g psi) = g (Q[U] psi) with U^g_mu(x) = g(x) U_mu(x) g(x+mu)^*."""
    rng = np.random.default_rng(9)
    L1, L2 = 8, 6
    U = random_links(L1, L2, rng, unitary=True)
    psi = random_spinor(L1, L2, rng)
    g = np.exp(1j * rng.uniform(-np.pi, np.pi, (L2, L1)))
    Ug = np.empty_like(U)
    Ug[0] = g * U[0] * np.conj(np.roll(g, -1, axis=1))
    Ug[1] = g * U[1] * np.conj(np.roll(g, -1, axis=0))
    lhs = O.gamm5Dw(Ug, g[None] * psi, 0.25)
    rhs = g[None] * O.gamm5Dw(U, psi, 0.25)
    assert np.max(np.abs(lhs - rhs)) < 1e-13 * np.max(np.abs(rhs))


def test_antiperiodic_equals_twisted_links(O):
    """The antiperiodic flag equals flipping the sign of U2 on the last row."""
    rng = np.random.default_rng(2)
    L1, L2 = 6, 6
    U = random_links(L1, L2, rng)
    si = random_spinor(L1, L2, rng)
    Ut = U.copy()
    Ut[1, L2 - 1, :] *= -1
    np.testing.assert_allclose(O.gamm5Dw(U, si, 0.4, 1), O.gamm5Dw(Ut, si, 0.4, 0), rtol=0,
                               atol=1e-15)
