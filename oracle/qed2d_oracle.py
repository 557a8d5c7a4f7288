"""ctypes bindings of the FP64 CPU oracle (liboracle_qed2d.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, never by the product path.

Field shapes (numpy, C order) follow the reference's Julia column-major
[L1, L2, 2] layout:  spinors and links are complex128 arrays of shape
(2, L2, L1); momenta and forces float64 (2, L2, L1); per-site densities
float64 (L2, L1).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_qed2d.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


lib = _load()
_d = C.c_double
_i = C.c_int
_u32 = C.c_uint32
_u64 = C.c_uint64
_p = C.c_void_p
_pd = C.POINTER(C.c_double)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("orc_philox4x32_10", None, [C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)])
_sig("orc_det_log", _d, [_d])
_sig("orc_det_sincos2pi", None, [_d, _pd, _pd])
_sig("orc_fill_momenta", None, [_p, _i, _i, _u64, _u32, _u32])
_sig("orc_fill_pseudofermion", None, [_p, _i, _i, _u64, _u32, _u32, _u32])
_sig("orc_metropolis_uniform", _d, [_u64, _u32, _u32])
_sig("orc_hot_start", None, [_p, _i, _i, _u64, _u32])
_sig("orc_gamm5Dw", None, [_p, _p, _p, _d, _i, _i, _i])
_sig("orc_gamm5", None, [_p, _p, _i, _i])
_sig("orc_Dw", None, [_p, _p, _p, _d, _i, _i, _i])
_sig("orc_gamm5Dw_sqr", None, [_p, _p, _p, _d, _i, _i, _i])
_sig("orc_gamm5Dw_sqr_msq", None, [_p, _p, _p, _p, _d, _i, _i, _i])
_sig("orc_gamm5Dw_sqr_musq", None, [_p, _p, _p, _p, _d, _d, _i, _i, _i])
_sig("orc_gamm5Dw_sqr_sqr", None, [_p, _p, _p, _d, _i, _i, _i])
_sig("orc_gamm5Dw_sqr_sqr_musq", None, [_p, _p, _p, _p, _d, _d, _i, _i, _i])
_sig("orc_tr_dQwdU", None, [_p, _p, _p, _p, _i, _i, _i])
_sig("orc_krnl_force", None, [_p, _p, _p, _d, _i, _i])
_sig("orc_updt_fld", None, [_p, _p, _d, _i, _i])
_sig("orc_plaquette_field", None, [_p, _p, _i, _i])
_sig("orc_qtop_field", None, [_p, _p, _i, _i])
_sig("orc_action", _d, [_p, _d, _i, _i])
_sig("orc_plaquette", _d, [_p, _i, _i])
_sig("orc_qtop", _d, [_p, _i, _i])
_sig("orc_hamiltonian", _d, [_p, _p, _d, _i, _i])
_sig("orc_last_dh", _d, [])
_sig("orc_omp_threads", _i, [])
_sig("orc_dot", None, [_p, _p, C.c_long, _pd, _pd])
_sig("orc_norm2", _d, [_p, C.c_long])
_sig("orc_cg", _i, [_p, _p, _p, _d, _d, _i, _i, _d, _i, _i, _i, _i, C.POINTER(_i)])
_sig("orc_multicg", _i, [_p, _p, _p, _d, _i, _p, _p, _i, _d, _i, _i, _i])
_sig("orc_R", _i, [_p, _p, _p, _d, _i, _p, _p, _d, _d, _i, _d, _i, _i, _i])
_sig("orc_generate_pseudofermion", _i, [_p, _p, _p, _d, _i, _p, _p, _i, _d, _i, _i, _i])
_sig("orc_power_method", None, [_p, _p, _d, _i, _i, _i, _i, _pd, _pd])
_sig("orc_ellipk", _d, [_d])
_sig("orc_ellip_sncn", None, [_d, _d, _pd, _pd])
_sig("orc_zolotarev", None, [_i, _d, _d, _pd, _pd, _pd, _p, _p, _p])
_sig("orc_unfold", None, [_p, _p, _i, _i, _i])
_sig("orc_fill_cgauss", None, [_p, _i, _i, _u64, _u32, _u32, _u32, _u32])
_sig("orc_luscher_z", _i, [_p, _p, _p, _d, _i, _p, _p, _d, _d, _i, _d, _i, _i, _i])
_sig("orc_luscher_zp", _i, [_p, _i, _p, _p, _d, _i, _p, _p, _d, _d, _i, _d, _i, _i, _i])
_sig("orc_reweighting_factor", _d, [_p, _d, _i, _p, _p, _d, _d, _i, _i, _i, _d, _u64, _u32, _u32,
                                    _i, _i, _i, C.POINTER(_i)])


class Latt(C.Structure):
    _fields_ = [("L1", _i), ("L2", _i), ("bc", _i), ("beta", _d)]


_sig("orc_hmc_pure", _i, [_p, _d, _i, _i, _i, _u64, _u32, _u32, C.POINTER(Latt), _pd, _pd,
                          C.POINTER(_i)])
_sig("orc_hmc_nf2", _i, [_p, _d, _d, _i, _i, _i, _d, _i, _u64, _u32, _u32, C.POINTER(Latt),
                         _pd, _pd, C.POINTER(_i), C.POINTER(_i)])
_sig("orc_hmc_rhmc", _i, [_p, _p, _i, _p, _p, _p, _p, _d, _i, _i, _i, _d, _i, _i, _u64, _u32,
                          _u32, C.POINTER(Latt), _pd, _pd, C.POINTER(_i)])
_sig("orc_update_momenta_rhmc", _i, [_p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _d, _i, _d,
                                     C.POINTER(Latt)])
_sig("orc_update_momenta_nf2", _i, [_p, _p, _p, _p, _p, _d, _d, _i, _d, C.POINTER(Latt)])
_sig("orc_omf4_pure", None, [_p, _p, _d, _i, C.POINTER(Latt)])
_sig("orc_leapfrog_pure", None, [_p, _p, _d, _i, C.POINTER(Latt)])
_sig("orc_leapfrog_nf2", _i, [_p, _p, _p, _p, _p, _d, _d, _i, _i, _d, C.POINTER(Latt)])
_sig("orc_omf4_nf2", _i, [_p, _p, _p, _p, _p, _d, _d, _i, _i, _d, C.POINTER(Latt)])


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _shape(a):
    return a.shape[-1], a.shape[-2]          # (L1, L2)


def _cspinor(L1, L2):
    return np.zeros((2, L2, L1), np.complex128)


# ---- random numbers -------------------------------------------------------
def philox(ctr, key):
    c = (_u32 * 4)(*ctr)
    k = (_u32 * 2)(*key)
    o = (_u32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def det_log(u):
    return lib.orc_det_log(u)


def det_sincos2pi(u):
    s, c = _d(), _d()
    lib.orc_det_sincos2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


def momenta(L1, L2, seed, chain=0, traj=0):
    m = np.zeros((2, L2, L1))
    lib.orc_fill_momenta(_ptr(m), L1, L2, seed, chain, traj)
    return m


def pseudofermion(L1, L2, seed, chain=0, traj=0, flavour=0):
    x = _cspinor(L1, L2)
    lib.orc_fill_pseudofermion(_ptr(x), L1, L2, seed, chain, traj, flavour)
    return x


def cgauss(L1, L2, seed, chain, traj, purpose, sub0=0):
    x = _cspinor(L1, L2)
    lib.orc_fill_cgauss(_ptr(x), L1, L2, seed, chain, traj, purpose, sub0)
    return x


def metropolis_uniform(seed, chain=0, traj=0):
    return lib.orc_metropolis_uniform(seed, chain, traj)


def hot_start(L1, L2, seed, chain=0):
    U = _cspinor(L1, L2)
    lib.orc_hot_start(_ptr(U), L1, L2, seed, chain)
    return U


# ---- operators ------------------------------------------------------------
def gamm5Dw(U, si, am0, bc=0):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    lib.orc_gamm5Dw(_ptr(so), _ptr(U), _ptr(si), am0, L1, L2, bc)
    return so


def gamm5(si):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    lib.orc_gamm5(_ptr(so), _ptr(si), L1, L2)
    return so


def Dw(U, si, am0, bc=0):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    lib.orc_Dw(_ptr(so), _ptr(U), _ptr(si), am0, L1, L2, bc)
    return so


def gamm5Dw_sqr(U, si, am0, bc=0):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    lib.orc_gamm5Dw_sqr(_ptr(so), _ptr(U), _ptr(si), am0, L1, L2, bc)
    return so


def gamm5Dw_sqr_musq(U, si, am0, mu, bc=0):
    L1, L2 = _shape(si)
    so, tmp = _cspinor(L1, L2), _cspinor(L1, L2)
    lib.orc_gamm5Dw_sqr_musq(_ptr(so), _ptr(tmp), _ptr(U), _ptr(si), am0, mu, L1, L2, bc)
    return so


def gamm5Dw_sqr_sqr(U, si, am0, bc=0):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    lib.orc_gamm5Dw_sqr_sqr(_ptr(so), _ptr(U), _ptr(si), am0, L1, L2, bc)
    return so


def gamm5Dw_sqr_sqr_musq(U, si, am0, mu, bc=0):
    L1, L2 = _shape(si)
    so, tmp = _cspinor(L1, L2), _cspinor(L1, L2)
    lib.orc_gamm5Dw_sqr_sqr_musq(_ptr(so), _ptr(tmp), _ptr(U), _ptr(si), am0, mu, L1, L2, bc)
    return so


def tr_dQwdU(U, X, Y, bc=0):
    L1, L2 = _shape(X)
    frc = np.zeros((2, L2, L1))
    lib.orc_tr_dQwdU(_ptr(frc), _ptr(U), _ptr(X), _ptr(Y), L1, L2, bc)
    return frc


def krnl_force(U, beta):
    L1, L2 = _shape(U)
    f1, f2 = np.zeros((2, L2, L1)), np.zeros((2, L2, L1))
    lib.orc_krnl_force(_ptr(f1), _ptr(f2), _ptr(U), beta, L1, L2)
    return f1, f2


def updt_fld(U, mom, eps):
    L1, L2 = _shape(U)
    U = np.ascontiguousarray(U.copy())
    lib.orc_updt_fld(_ptr(U), _ptr(np.ascontiguousarray(mom)), eps, L1, L2)
    return U


def plaquette_field(U):
    L1, L2 = _shape(U)
    p = np.zeros((L2, L1))
    lib.orc_plaquette_field(_ptr(p), _ptr(U), L1, L2)
    return p


def qtop_field(U):
    L1, L2 = _shape(U)
    p = np.zeros((L2, L1))
    lib.orc_qtop_field(_ptr(p), _ptr(U), L1, L2)
    return p


def action(U, beta):
    L1, L2 = _shape(U)
    return lib.orc_action(_ptr(U), beta, L1, L2)


def plaquette(U):
    L1, L2 = _shape(U)
    return lib.orc_plaquette(_ptr(U), L1, L2)


def qtop(U):
    L1, L2 = _shape(U)
    return lib.orc_qtop(_ptr(U), L1, L2)


def hamiltonian(mom, U, beta):
    L1, L2 = _shape(U)
    return lib.orc_hamiltonian(_ptr(mom), _ptr(U), beta, L1, L2)


def dot(a, b):
    re, im = _d(), _d()
    lib.orc_dot(_ptr(a), _ptr(b), a.size, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def norm2(a):
    return lib.orc_norm2(_ptr(a), a.size)


# ---- solvers --------------------------------------------------------------
def cg(U, si, am0, maxiter, eps, mu=0.0, op=0, x0=None, bc=0):
    """Returns (x, iterations, status)."""
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2) if x0 is None else np.ascontiguousarray(x0.copy())
    st = _i()
    it = lib.orc_cg(_ptr(so), _ptr(U), _ptr(si), am0, mu, op, maxiter, eps,
                    0 if x0 is None else 1, L1, L2, bc, C.byref(st))
    return so, it, st.value


def multicg(U, si, am0, mu, rho, maxiter, eps, bc=0):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    mu = np.ascontiguousarray(mu, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    bad = lib.orc_multicg(_ptr(so), _ptr(U), _ptr(si), am0, len(mu), _ptr(mu), _ptr(rho),
                          maxiter, eps, L1, L2, bc)
    return so, bad


def R(U, si, am0, z, maxiter, eps, bc=0):
    L1, L2 = _shape(si)
    so = _cspinor(L1, L2)
    mu = np.ascontiguousarray(z["mu"], np.float64)
    rho = np.ascontiguousarray(z["rho"], np.float64)
    bad = lib.orc_R(_ptr(so), _ptr(U), _ptr(si), am0, len(mu), _ptr(mu), _ptr(rho), z["A"],
                    z["r_b"], maxiter, eps, L1, L2, bc)
    return so, bad


def generate_pseudofermion(U, X, am0, z, maxiter, tol, bc=0):
    L1, L2 = _shape(X)
    F = _cspinor(L1, L2)
    mu = np.ascontiguousarray(z["mu"], np.float64)
    nu = np.ascontiguousarray(z["nu"], np.float64)
    bad = lib.orc_generate_pseudofermion(_ptr(F), _ptr(U), _ptr(X), am0, len(mu), _ptr(mu),
                                         _ptr(nu), maxiter, tol, L1, L2, bc)
    return F, bad


def power_method(U, b0, am0, iters, bc=0):
    L1, L2 = _shape(U)
    lmin, lmax = _d(), _d()
    lib.orc_power_method(_ptr(U), _ptr(b0), am0, iters, L1, L2, bc, C.byref(lmin),
                         C.byref(lmax))
    return lmin.value, lmax.value


def ellipk(m):
    return lib.orc_ellipk(m)


def ellip_sncn(u, m):
    s, c = _d(), _d()
    lib.orc_ellip_sncn(u, m, C.byref(s), C.byref(c))
    return s.value, c.value


def zolotarev(n, r_a, r_b):
    eps, A, delta = _d(), _d(), _d()
    rho, mu, nu = np.zeros(n), np.zeros(n), np.zeros(n)
    lib.orc_zolotarev(n, r_a, r_b, C.byref(eps), C.byref(A), C.byref(delta), _ptr(rho),
                      _ptr(mu), _ptr(nu))
    return {"n": n, "r_a": r_a, "r_b": r_b, "eps": eps.value, "A": A.value,
            "delta": delta.value, "rho": rho, "mu": mu, "nu": nu}


# ---- MD / HMC -------------------------------------------------------------
def latt(L1, L2, beta, bc=0):
    return Latt(L1, L2, bc, beta)


def omf4_pure(mom, U, eps, ns, p):
    lib.orc_omf4_pure(_ptr(mom), _ptr(U), eps, ns, C.byref(p))


def leapfrog_pure(mom, U, eps, ns, p):
    lib.orc_leapfrog_pure(_ptr(mom), _ptr(U), eps, ns, C.byref(p))


def leapfrog_nf2(mom, U, X, F, am0, eps, ns, maxiter, tol, p):
    g5DX = np.zeros_like(X)
    return lib.orc_leapfrog_nf2(_ptr(mom), _ptr(U), _ptr(X), _ptr(F), _ptr(g5DX), am0, eps, ns,
                                maxiter, tol, C.byref(p))


def omf4_nf2(mom, U, X, F, am0, eps, ns, maxiter, tol, p):
    g5DX = np.zeros_like(X)
    return lib.orc_omf4_nf2(_ptr(mom), _ptr(U), _ptr(X), _ptr(F), _ptr(g5DX), am0, eps, ns,
                            maxiter, tol, C.byref(p))


def hmc_pure(U, eps, ns, p, seed, traj, chain=0, integrator=1, qzero=0):
    hi, hf, acc = _d(), _d(), _i()
    lib.orc_hmc_pure(_ptr(U), eps, ns, integrator, qzero, seed, chain, traj, C.byref(p),
                     C.byref(hi), C.byref(hf), C.byref(acc))
    return hi.value, hf.value, acc.value


def omp_threads():
    """OpenMP threads of the oracle's parallel regions."""
    return lib.orc_omp_threads()


def last_dh():
    """dH of the last hmc_nf2 trajectory from extended-precision energies."""
    return lib.orc_last_dh()


def hmc_nf2(U, am0, eps, ns, maxiter, tol, p, seed, traj, chain=0, integrator=0, qzero=0):
    hi, hf, acc, it = _d(), _d(), _i(), _i()
    rc = lib.orc_hmc_nf2(_ptr(U), am0, eps, ns, integrator, maxiter, tol, qzero, seed, chain,
                         traj, C.byref(p), C.byref(hi), C.byref(hf), C.byref(acc), C.byref(it))
    return hi.value, hf.value, acc.value, it.value, rc


def hmc_rhmc(U, am0s, zs, eps, ns, maxiter, tol, p, seed, traj, chain=0, qzero=0,
             reversibility_test=0, integrator=0):
    nf = len(am0s)
    am0 = np.ascontiguousarray(am0s, np.float64)
    npole = np.ascontiguousarray([z["n"] for z in zs], np.int32)
    keep = [np.ascontiguousarray(z[k], np.float64) for z in zs for k in ("mu", "nu", "rho")]
    mus = (C.c_void_p * nf)(*[keep[3 * j].ctypes.data for j in range(nf)])
    nus = (C.c_void_p * nf)(*[keep[3 * j + 1].ctypes.data for j in range(nf)])
    rhos = (C.c_void_p * nf)(*[keep[3 * j + 2].ctypes.data for j in range(nf)])
    hi, hf, acc = _d(), _d(), _i()
    rc = lib.orc_hmc_rhmc(_ptr(U), _ptr(am0), nf, _ptr(npole), mus, nus, rhos, eps, ns,
                          integrator, maxiter, tol, qzero, reversibility_test, seed, chain, traj,
                          C.byref(p),
                          C.byref(hi), C.byref(hf), C.byref(acc))
    return hi.value, hf.value, acc.value, rc


def luscher_z(U, X, am0, z, maxiter, tol, p=1, bc=0):
    """Z^p X with Z = Q^2 R^2 - 1."""
    L1, L2 = _shape(X)
    Z = _cspinor(L1, L2)
    mu = np.ascontiguousarray(z["mu"], np.float64)
    rho = np.ascontiguousarray(z["rho"], np.float64)
    bad = lib.orc_luscher_zp(_ptr(Z), p, _ptr(U), _ptr(np.ascontiguousarray(X)), am0, len(mu),
                             _ptr(mu), _ptr(rho), z["A"], z["r_b"], maxiter, tol, L1, L2, bc)
    return Z, bad


def reweighting_factor(U, am0, z, N, taylor, maxiter, tol, seed, chain=0, traj=0, bc=0):
    L1, L2 = _shape(U)
    mu = np.ascontiguousarray(z["mu"], np.float64)
    rho = np.ascontiguousarray(z["rho"], np.float64)
    st = _i()
    W = lib.orc_reweighting_factor(_ptr(U), am0, len(mu), _ptr(mu), _ptr(rho), z["A"], z["r_b"],
                                   N, taylor, maxiter, tol, seed, chain, traj, L1, L2, bc,
                                   C.byref(st))
    return W, st.value


def unfold(U, reflect=False):
    L1, L2 = _shape(U)
    out = np.zeros((2, 2 * L2, 2 * L1), np.complex128)
    lib.orc_unfold(_ptr(out), _ptr(U), L1, L2, 1 if reflect else 0)
    return out
