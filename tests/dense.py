"""Dense-matrix brute force for small lattices, built independently of the
stencil code from the Dirac-matrix form of the Wilson operator

    Dw = (2 + m) - 1/2 sum_mu [ (1 - g_mu) U_mu(x) d_{y,x+mu}
                               + (1 + g_mu) U_mu(x-mu)^* d_{y,x-mu} ],
    Q  = g5 Dw,   g1 = sigma1, g2 = sigma2, g5 = sigma3

(the gamma-matrix reading of reference src/QED2dDirac.jl:2-33).  Vectors are
flattened from the (2, L2, L1) field layout.
"""
import numpy as np

S1 = np.array([[0, 1], [1, 0]], complex)
S2 = np.array([[0, -1j], [1j, 0]], complex)
S3 = np.array([[1, 0], [0, -1]], complex)
I2 = np.eye(2, dtype=complex)


def _idx(c, i1, i2, L1, L2):
    return c * L1 * L2 + i2 * L1 + i1


def dirac_dense(U, am0, bc=0):
    """Dw as a dense (2V x 2V) matrix.  U: complex (2, L2, L1)."""
    _, L2, L1 = U.shape
    V = L1 * L2
    D = np.zeros((2 * V, 2 * V), complex)
    gam = [S1, S2]
    for i2 in range(L2):
        for i1 in range(L1):
            for a in range(2):
                D[_idx(a, i1, i2, L1, L2), _idx(a, i1, i2, L1, L2)] += 2.0 + am0
            for mu in range(2):
                if mu == 0:
                    f = ((i1 + 1) % L1, i2)
                    b = ((i1 - 1) % L1, i2)
                    uf = U[0, i2, i1]
                    ub = np.conj(U[0, i2, (i1 - 1) % L1])
                    sf = sb = 1.0
                else:
                    f = (i1, (i2 + 1) % L2)
                    b = (i1, (i2 - 1) % L2)
                    uf = U[1, i2, i1]
                    ub = np.conj(U[1, (i2 - 1) % L2, i1])
                    sf = -1.0 if (bc & 1 and i2 == L2 - 1) else 1.0
                    sb = -1.0 if (bc & 1 and i2 == 0) else 1.0
                Pm = 0.5 * (I2 - gam[mu])
                Pp = 0.5 * (I2 + gam[mu])
                for a in range(2):
                    for c in range(2):
                        D[_idx(a, i1, i2, L1, L2), _idx(c, f[0], f[1], L1, L2)] -= sf * Pm[a, c] * uf
                        D[_idx(a, i1, i2, L1, L2), _idx(c, b[0], b[1], L1, L2)] -= sb * Pp[a, c] * ub
    return D


def gamma5_dense(V):
    g = np.zeros((2 * V, 2 * V), complex)
    g[:V, :V] = np.eye(V)
    g[V:, V:] = -np.eye(V)
    return g


def q_dense(U, am0, bc=0):
    _, L2, L1 = U.shape
    return gamma5_dense(L1 * L2) @ dirac_dense(U, am0, bc)


def vec(f):
    return np.ascontiguousarray(f).reshape(-1)


def field(v, L1, L2):
    return np.ascontiguousarray(v.reshape(2, L2, L1))


def random_links(L1, L2, rng, unitary=False):
    """Non-unitary links U = (N + iN)/sqrt(2) as the reference tests use
    (test/rhmc_pf_forcetest.jl:20), or random U(1) phases."""
    if unitary:
        th = rng.uniform(-np.pi, np.pi, (2, L2, L1))
        return np.exp(1j * th)
    return (rng.standard_normal((2, L2, L1)) + 1j * rng.standard_normal((2, L2, L1))) / np.sqrt(2)


def random_spinor(L1, L2, rng):
    return (rng.standard_normal((2, L2, L1)) + 1j * rng.standard_normal((2, L2, L1))) / np.sqrt(2)
