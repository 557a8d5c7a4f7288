"""GPU parity of the solvers (CG v1/v2, mixed precision, MultiCG, R,
pseudofermion generation, power method) and of the MD / HMC drivers against
the FP64 CPU oracle.

Tolerances: solutions are compared through the true FP64 residual and to the
oracle's solution at the level the stopping rule allows; HMC energies agree to
1e-9 absolute at 24^2 and the accept/reject sequence is identical."""
import numpy as np
import pytest

from dense import random_links, random_spinor
from gpu_util import dev, host, relerr, say

pytestmark = pytest.mark.gpu


def _true_res(O, U, x, b, am0, mu=0.0, bc=0):
    Ax = O.gamm5Dw_sqr_musq(U, x, am0, mu, bc)
    return O.norm2(b - Ax) / O.norm2(b)


@pytest.mark.parametrize("L1,L2", [(24, 24), (64, 40), (33, 20)])
@pytest.mark.parametrize("prec", [0, 1])
def test_cg_v1(Q, O, L1, L2, prec):
    say("CG v1", L1, L2, "prec", prec)
    rng = np.random.default_rng(L1)
    U = random_links(L1, L2, rng, unitary=True)
    b = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    x = Q.spinor(prm)
    tmp = Q.spinor(prm)
    it = Q.CG(x, dev(U), dev(b), 0.2, 1000, 1e-20, Q.gamm5Dw_sqr_msq, prm, Q.KernelParm(),
              precision=prec)
    xr, itr, st = O.cg(U, b, 0.2, 1000, 1e-20)
    assert st == 0
    assert _true_res(O, U, host(x), b, 0.2) < 1e-20
    assert relerr(host(x), xr) < 1e-8
    if prec == 0:
        assert abs(it - itr) <= 1
    else:
        assert it <= 1.3 * itr + 10


def test_cg_v2_shift_and_guess(Q, O):
    say("CG v2")
    L1, L2 = 32, 32
    rng = np.random.default_rng(4)
    U = random_links(L1, L2, rng)
    b = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    x = Q.spinor(prm)
    dU, db = dev(U), dev(b)
    it = Q.CG(x, dU, db, 0.5, 0.6, 2000, 1e-22, Q.gamm5Dw_sqr_musq, prm, Q.KernelParm())
    xr, itr, st = O.cg(U, b, 0.5, 2000, 1e-22, mu=0.6, op=1)
    assert abs(it - itr) <= 1
    assert relerr(host(x), xr) < 1e-9
    # warm start from the converged solution: reference returns 0 iterations
    it2 = Q.CG(x, dU, db, 0.5, 0.6, 2000, 1e-22, Q.gamm5Dw_sqr_musq, prm, Q.KernelParm(),
               so_as_guess=True)
    assert it2 == 0
    # warm start from a different shift's solution converges to the new shift
    it3 = Q.CG(x, dU, db, 0.5, 0.9, 2000, 1e-22, Q.gamm5Dw_sqr_musq, prm, Q.KernelParm(),
               so_as_guess=True)
    xr3, _, _ = O.cg(U, b, 0.5, 2000, 1e-22, mu=0.9, op=1)
    assert it3 > 0 and relerr(host(x), xr3) < 1e-9
    # mixed precision with a guess
    x.zero_()
    Q.CG(x, dU, db, 0.5, 0.6, 2000, 1e-22, Q.gamm5Dw_sqr_musq, prm, Q.KernelParm(),
         precision=Q.PREC_MIXED)
    assert _true_res(O, U, host(x), b, 0.5, 0.6) < 1e-22


@pytest.mark.parametrize("rows", [1, 2, 8, 16])
def test_cg_tiles(Q, O, rows):
    """CG (FP64 and mixed) gives the oracle's solution for every tile height."""
    say("CG tiles", rows)
    L1, L2 = 48, 36
    rng = np.random.default_rng(rows)
    U = random_links(L1, L2, rng, unitary=True)
    b = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    Q.lib.qed2d_ctx_set_tile(prm.ctx(), rows)
    Q.lib.qed2d_ctx_set_tile_f32(prm.ctx(), rows)
    xr, itr, st = O.cg(U, b, 0.3, 2000, 1e-20)
    for prec in (0, 1):
        x = Q.spinor(prm)
        it = Q.CG(x, dev(U), dev(b), 0.3, 2000, 1e-20, Q.gamm5Dw_sqr_msq, prm, None,
                  precision=prec)
        assert _true_res(O, U, host(x), b, 0.3) < 1e-20
        assert relerr(host(x), xr) < 1e-8
        if prec == 0:
            assert abs(it - itr) <= 1


@pytest.mark.parametrize("k1,k2", [(2, 16), (16, 2), (4, 8)])
def test_cg_split_tiles(Q, O, k1, k2):
    """The two kernels of a CG iteration at different tile heights (each
    producer's partial-sum count travels with it): the oracle's solution,
    both precisions, odd L1 (one-site FP32 march) and even."""
    for L1, L2 in ((48, 36), (33, 20)):
        say("CG split tiles", k1, k2, L1, L2)
        rng = np.random.default_rng(100 * k1 + k2 + L1)
        U = random_links(L1, L2, rng, unitary=True)
        b = random_spinor(L1, L2, rng)
        prm = Q.LattParm((L1, L2), 4.0)
        Q.set_tiles(prm, cg64=k1, cg32=k1, cg64_k2=k2, cg32_k2=k2)
        t = Q.get_tiles(prm)
        assert (t["cg64"], t["cg32"], t["cg64_k2"], t["cg32_k2"]) == (k1, k1, k2, k2)
        xr, itr, st = O.cg(U, b, 0.3, 2000, 1e-20)
        for prec in (0, 1):
            x = Q.spinor(prm)
            it = Q.CG(x, dev(U), dev(b), 0.3, 2000, 1e-20, Q.gamm5Dw_sqr_msq, prm, None,
                      precision=prec)
            assert _true_res(O, U, host(x), b, 0.3) < 1e-20
            assert relerr(host(x), xr) < 1e-8
            if prec == 0:
                assert abs(it - itr) <= 1
        # set_tiles with a solver height resets the second kernel's to follow it
        Q.set_tiles(prm, cg64=8, cg32=8)
        t = Q.get_tiles(prm)
        assert (t["cg64_k2"], t["cg32_k2"]) == (8, 8)


@pytest.mark.parametrize("L1,L2,k2,mu,guess", [(64, 48, 0, 0.0, False), (33, 20, 0, 0.0, False),
                                               (64, 48, 2, 0.7, True),
                                               (2048, 512, 1, 0.0, False)])
def test_cg_mixed_reliable_modes(Q, O, L1, L2, k2, mu, guess):
    """Mixed-precision CG with reliable updates decided on the device (the
    default: requested by an FP32 iteration, run by kernels queued at chunk
    starts) and by the host after a pause: both reach the FP64 true residual,
    agree with the FP64 solve at the level the tolerance allows, and take
    about as many iterations.  Covers the site-pair kernels, the one-site FP32
    march (odd L1), a shifted operator with a guess, and a K2 grid past 1024
    blocks whose partial sums are folded (2048 x 512 at one row per thread)."""
    say("mixed CG reliable modes", L1, L2, k2, mu, guess)
    rng = np.random.default_rng(L1 + L2 + k2)
    U = random_links(L1, L2, rng, unitary=True)
    b = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    if k2:
        Q.set_tiles(prm, cg32=k2, cg32_k2=k2)
    dU, db = dev(U), dev(b)
    op = Q.gamm5Dw_sqr_musq if mu else Q.gamm5Dw_sqr_msq
    x0 = Q.spinor(prm)
    it0 = Q.CG(x0, dU, db, 0.3, mu, 20000, 1e-24, op, prm, None)
    res = {}
    for mode in (True, False):
        Q.set_reliable_mode(prm, mode)
        x = Q.spinor(prm)
        if guess:  # a guess half way: x0 / 2
            x.copy_(0.5 * x0)
        it = Q.CG(x, dU, db, 0.3, mu, 20000, 1e-24, op, prm, None, so_as_guess=guess,
                  precision=Q.PREC_MIXED)
        res[mode] = (it, x)
        if L1 * L2 <= 64 * 64:
            assert _true_res(O, U, host(x), b, 0.3, mu) < 1e-24
        assert relerr(host(x), host(x0)) < 1e-9, mode
    Q.set_reliable_mode(prm, True)
    say("iterations fp64", it0, "on device", res[True][0], "host", res[False][0])
    assert abs(res[True][0] - res[False][0]) <= 4
    assert res[True][0] <= 1.3 * it0 + 10


def test_cg_not_converged_raises(Q):
    say("CG maxiter")
    L1, L2 = 24, 24
    rng = np.random.default_rng(8)
    prm = Q.LattParm((L1, L2), 4.0)
    x = Q.spinor(prm)
    with pytest.raises(Q.CGNotConverged):
        Q.CG(x, dev(random_links(L1, L2, rng)), dev(random_spinor(L1, L2, rng)), 0.01, 3, 1e-30,
             Q.gamm5Dw_sqr_msq, prm, Q.KernelParm())


def test_cg_zero_rhs(Q):
    prm = Q.LattParm((16, 16), 4.0)
    x = Q.spinor(prm) + 1.0
    it = Q.CG(x, Q.cold_links(prm), Q.spinor(prm), 0.1, 10, 1e-12, Q.gamm5Dw_sqr_msq, prm,
              Q.KernelParm())
    assert it == 0 and float(x.abs().max()) == 0.0


def test_multicg_R_pseudofermion(Q, O):
    say("MultiCG / R / generate_pseudofermion")
    L1, L2 = 16, 16
    rng = np.random.default_rng(12)
    U = random_links(L1, L2, rng)
    X = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 6.05)
    rp = Q.get_rhmc_params(4, np.sqrt(0.1), np.sqrt(26.0))
    z = O.zolotarev(4, np.sqrt(0.1), np.sqrt(26.0))
    assert relerr(rp.rho, z["rho"]) < 1e-14 and relerr(rp.mu, z["mu"]) < 1e-14
    dU, dX = dev(U), dev(X)
    so = Q.spinor(prm)
    Q.MultiCG(so, dU, dX, 0.3, 5000, 1e-24, Q.gamm5Dw_sqr_musq, rp, prm)
    ref, bad = O.multicg(U, X, 0.3, z["mu"], z["rho"], 5000, 1e-24)
    assert relerr(host(so), ref) < 1e-10
    Q.R(so, dU, dX, 0.3, 5000, 1e-24, Q.gamm5Dw_sqr_musq, rp, prm)
    ref, bad = O.R(U, X, 0.3, z, 5000, 1e-24)
    assert relerr(host(so), ref) < 1e-10
    F = Q.spinor(prm)
    Q.generate_pseudofermion(F, dU, dX, 0.3, 5000, 1e-24, prm, None, rp)
    Fr, bad = O.generate_pseudofermion(U, X, 0.3, z, 5000, 1e-24)
    assert relerr(host(F), Fr) < 1e-10
    # test/rhmc_pf_generation.jl:56-57: X^dag X = phi^dag MultiCG(phi)
    Q.MultiCG(so, dU, F, 0.3, 5000, 1e-24, Q.gamm5Dw_sqr_musq, rp, prm)
    assert abs(Q.dot(so, F, prm).real - O.norm2(X)) < 1e-10 * O.norm2(X)


@pytest.mark.parametrize("L,am0,iters", [(8, 0.5, 2000), (32, 0.5, 2000), (24, 0.2, 5000)])
def test_power_method(Q, O, L, am0, iters):
    """power_method (src/QED2dRHMC.jl:135-170).  After the lambda_max loop the
    low modes survive only as rounding noise, so the lambda_min loop amplifies
    device-specific rounding: GPU and oracle are each checked against the dense
    spectrum on cases where the iteration converges (the oracle to 4e-15 of
    lambda_max at 32^2 and 24^2), and against each other for lambda_max."""
    from dense import q_dense
    say("power method", L, am0, iters)
    L1, L2 = L, L
    rng = np.random.default_rng(21)
    U = random_links(L1, L2, rng)
    b0 = random_spinor(L1, L2, rng)
    ev = np.linalg.eigvalsh(q_dense(U, am0) @ q_dense(U, am0))
    prm = Q.LattParm((L1, L2), 6.05)
    lmin, lmax = Q.power_method(dev(U), am0, prm, iter=iters, b0=dev(b0))
    rmin, rmax = O.power_method(U, b0, am0, iters)
    assert abs(lmax - rmax) < 1e-12 * rmax
    assert abs(lmax - ev[-1]) < 1e-10 * ev[-1]
    assert abs(lmin - ev[0]) < 1e-8 * ev[-1]
    assert abs(rmin - ev[0]) < 1e-8 * ev[-1]


def test_omf4_pure_gauge(Q, O):
    say("OMF4 pure gauge")
    L1, L2 = 32, 32
    U = O.hot_start(L1, L2, 77)
    mom = O.momenta(L1, L2, 77, 0, 0)
    prm = Q.LattParm((L1, L2), 4.0)
    dU, dm = dev(U), dev(mom)
    Q.OMF4_pure(dm, dU, 0.1, 10, prm)
    p = O.latt(L1, L2, 4.0)
    O.omf4_pure(mom, U, 0.1, 10, p)
    assert relerr(host(dU), U) < 1e-12
    assert relerr(host(dm), mom) < 1e-11


@pytest.mark.parametrize("L1,L2", [(33, 20), (64, 7), (128, 96)])
def test_pure_gauge_integrators_shapes(Q, O, L1, L2):
    """The pure-gauge integrators (each link update fused with the momentum
    update after it, out of place between two buffer pairs): OMF4 for 1, 2 and 3
    steps (an odd number of fused passes ends in the scratch pair and is copied
    back) and pure-gauge HMC with leapfrog and OMF4, against the oracle, on
    shapes with wave edges, odd widths and rows the tiles do not divide."""
    say("pure gauge integrators", L1, L2)
    prm = Q.LattParm((L1, L2), 3.0)
    p = O.latt(L1, L2, 3.0)
    for ns in (1, 2, 3):
        U = O.hot_start(L1, L2, 70 + ns)
        mom = O.momenta(L1, L2, 70 + ns, 0, 0)
        dU, dm = dev(U), dev(mom)
        Q.OMF4_pure(dm, dU, 0.07, ns, prm)
        O.omf4_pure(mom, U, 0.07, ns, p)
        assert relerr(host(dU), U) < 1e-12, ns
        assert relerr(host(dm), mom) < 1e-11, ns
    for code, integ in ((0, Q.Leapfrog()), (1, Q.OMF4())):
        U = O.hot_start(L1, L2, 90 + code)
        dU = dev(U)
        Q.seed(4242)
        for t in range(3):
            info = Q.HMC(dU, 0.05, 5, [], prm, None, integrator=integ)
            hi, hf, a = O.hmc_pure(U, 0.05, 5, p, 4242, t, integrator=code)
            assert abs(info.hini - hi) < 1e-9 and abs(info.hfin - hf) < 1e-9, (code, t)
            assert info.accepted == a, (code, t)
        assert relerr(host(dU), U) < 1e-10, code


@pytest.mark.parametrize("integ", [0, 1])
def test_md_nf2(Q, O, integ):
    say("MD nf2 integrator", integ)
    L1, L2 = 16, 16
    U = O.hot_start(L1, L2, 5)
    mom = O.momenta(L1, L2, 5, 0, 1)
    X = O.pseudofermion(L1, L2, 5, 0, 1)
    F = O.gamm5Dw(U, X, 0.5)
    prm = Q.LattParm((L1, L2), 4.0)
    dU, dm, dX, dF, dg = dev(U), dev(mom), Q.spinor(prm), dev(F), Q.spinor(prm)
    I = Q.OMF4() if integ else Q.Leapfrog()
    Q.MD_nf2(I, dm, dU, dX, dF, dg, 0.5, 0.05, 4, 5000, 1e-24, prm)
    p = O.latt(L1, L2, 4.0)
    Xo = np.zeros_like(X)
    if integ:
        O.omf4_nf2(mom, U, Xo, F, 0.5, 0.05, 4, 5000, 1e-24, p)
    else:
        O.leapfrog_nf2(mom, U, Xo, F, 0.5, 0.05, 4, 5000, 1e-24, p)
    assert relerr(host(dU), U) < 1e-10
    assert relerr(host(dm), mom) < 1e-9
    assert relerr(host(dX), Xo) < 1e-8


def test_hmc_pure_sequence(Q, O):
    say("HMC pure")
    L1, L2 = 24, 24
    prm = Q.LattParm((L1, L2), 4.0)
    U = O.hot_start(L1, L2, 1234)
    dU = dev(U)
    p = O.latt(L1, L2, 4.0)
    Q.seed(1234)
    acc = []
    for t in range(5):
        info = Q.HMC(dU, 0.1, 10, acc, prm, None)
        hi, hf, a = O.hmc_pure(U, 0.1, 10, p, 1234, t)
        assert abs(info.hini - hi) < 1e-9 and abs(info.hfin - hf) < 1e-9
        assert info.accepted == a
    assert relerr(host(dU), U) < 1e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_hmc_nf2_sequence(Q, O, prec):
    """Nf=2 at 24^2 (beta 4, m0 0.2, leapfrog ns=10 eps=0.1) with a
    tight CG tolerance, so that FP64 and mixed MD forces agree with the oracle
    to the level the energies are compared at."""
    say("HMC nf2 prec", prec)
    L1, L2 = 24, 24
    prm = Q.LattParm((L1, L2), 4.0)
    U = O.hot_start(L1, L2, 1234)
    dU = dev(U)
    p = O.latt(L1, L2, 4.0)
    Q.seed(1234)
    acc = []
    for t in range(4):
        info = Q.HMC(dU, 0.2, 0.1, 10, acc, 10000, 1e-24, prm, None, precision=prec)
        hi, hf, a, it, rc = O.hmc_nf2(U, 0.2, 0.1, 10, 10000, 1e-24, p, 1234, t)
        assert rc == 0
        assert abs(info.hini - hi) < 1e-9
        assert abs((info.hfin - info.hini) - (hf - hi)) < 1e-8
        assert info.accepted == a
    assert relerr(host(dU), U) < 1e-8


def test_hmc_rhmc_reversibility(Q, O):
    """test/rhmc_hmc_reversibility.jl:41-50 restated: forward + backward MD
    returns the gauge action to 1e-10."""
    say("RHMC reversibility")
    L1, L2 = 10, 10
    rng = np.random.default_rng(1234)
    U = random_links(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 6.05)
    dU = dev(U)
    b0 = dev(random_spinor(L1, L2, rng))
    lmin, lmax = Q.power_method(dU, 10.0, prm, iter=1000, b0=b0)
    rp = Q.get_rhmc_params([5], [np.sqrt(lmin)], [np.sqrt(lmax)])
    S0 = Q.Action(dU, prm)
    Q.seed(1234)
    info = Q.HMC(dU, [10.0], 0.01, 100, [], 10000, 1e-25, prm, None, rp,
                 reversibility_test=True)
    assert info.accepted == -1
    assert abs(Q.Action(dU, prm) - S0) < 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_hmc_rhmc_vs_oracle(Q, O, precision):
    """Two RHMC trajectories against the oracle; precision 1: the force's
    solves in mixed precision (FP32 multi-shift, then every pole by the
    batched mixed-precision CG from that guess), at the FP64 tolerances."""
    say("RHMC vs oracle, precision", precision)
    L1, L2 = 12, 12
    prm = Q.LattParm((L1, L2), 4.0)
    U = O.hot_start(L1, L2, 99)
    dU = dev(U)
    p = O.latt(L1, L2, 4.0)
    z = O.zolotarev(6, np.sqrt(0.05), np.sqrt(20.0))
    rp = Q.get_rhmc_params(6, np.sqrt(0.05), np.sqrt(20.0))
    Q.seed(99)
    for t in range(2):
        info = Q.HMC(dU, [0.5, 0.7], 0.1, 5, [], 10000, 1e-24, prm, None, [rp, rp],
                     precision=precision)
        hi, hf, a, rc = O.hmc_rhmc(U, [0.5, 0.7], [z, z], 0.1, 5, 10000, 1e-24, p, 99, t)
        assert rc == 0
        assert abs(info.hini - hi) < 1e-9
        assert abs(info.dH - (hf - hi)) < 1e-8
        assert info.accepted == a


@pytest.mark.parametrize("n", [1, 5, 16])
def test_multishift_cg_vs_sequential_oracle(Q, O, n):
    """Multi-shift CG against the oracle's per-shift CG (the reference's
    MultiCG loop, src/QED2dRHMC.jl:284-287): each x_i to the tolerance."""
    say("multishift", n)
    L1, L2 = 32, 24
    rng = np.random.default_rng(n)
    U = random_links(L1, L2, rng, unitary=True)
    b = random_spinor(L1, L2, rng)
    z = O.zolotarev(n, np.sqrt(0.01), np.sqrt(17.0))
    prm = Q.LattParm((L1, L2), 4.0)
    xs = [Q.spinor(prm) for _ in range(n)]
    it = Q.multishift_CG(xs, dev(U), dev(b), 0.15, list(z["mu"]), 5000, 1e-22, prm)
    assert it > 0
    for k in range(n):
        xr, itr, st = O.cg(U, b, 0.15, 5000, 1e-22, mu=z["mu"][k], op=1)
        assert _true_res(O, U, host(xs[k]), b, 0.15, z["mu"][k]) < 1e-21
        assert relerr(host(xs[k]), xr) < 1e-9


def test_multicg_multishift_and_sequential_agree(Q, O):
    say("MultiCG multishift vs sequential")
    L1, L2 = 24, 24
    rng = np.random.default_rng(77)
    U = random_links(L1, L2, rng)
    X = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    rp = Q.get_rhmc_params(8, np.sqrt(0.05), np.sqrt(30.0))
    z = O.zolotarev(8, np.sqrt(0.05), np.sqrt(30.0))
    dU, dX = dev(U), dev(X)
    a, b = Q.spinor(prm), Q.spinor(prm)
    Q.set_multishift(prm, True)
    Q.MultiCG(a, dU, dX, 0.3, 5000, 1e-24, Q.gamm5Dw_sqr_musq, rp, prm)
    Q.set_multishift(prm, False)
    Q.MultiCG(b, dU, dX, 0.3, 5000, 1e-24, Q.gamm5Dw_sqr_musq, rp, prm)
    Q.set_multishift(prm, True)
    ref, bad = O.multicg(U, X, 0.3, z["mu"], z["rho"], 5000, 1e-24)
    assert relerr(host(a), ref) < 1e-10 and relerr(host(b), ref) < 1e-10


@pytest.mark.parametrize("p", [1, 2])
def test_luscher_z_vs_oracle(Q, O, p):
    say("LuscherZp", p)
    L1, L2 = 24, 16
    rng = np.random.default_rng(31 + p)
    U = random_links(L1, L2, rng, unitary=True)
    X = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    rp = Q.get_rhmc_params(3, np.sqrt(0.02), np.sqrt(70.0))
    z = O.zolotarev(3, np.sqrt(0.02), np.sqrt(70.0))
    Z = Q.spinor(prm)
    Q.LuscherZp(Z, p, dev(U), dev(X), 0.25, 10000, 1e-26, rp, prm)
    ref, bad = O.luscher_z(U, X, 0.25, z, 10000, 1e-26, p=p)
    assert bad == 0
    # Z^p X = O(delta^p) X: the error is measured against X
    assert np.max(np.abs(host(Z) - ref)) / np.max(np.abs(X)) < 1e-11
    if p == 1:
        Z1 = Q.spinor(prm)
        Q.LuscherZ(Z1, dev(U), dev(X), 0.25, 10000, 1e-26, rp, prm)
        assert np.array_equal(host(Z1), host(Z))


def test_reweighting_factor_vs_oracle(Q, O):
    say("reweighting_factor")
    L1, L2 = 32, 24
    rng = np.random.default_rng(41)
    U = random_links(L1, L2, rng, unitary=True)
    prm = Q.LattParm((L1, L2), 4.0)
    rp = Q.get_rhmc_params(3, np.sqrt(0.02), np.sqrt(70.0), reweighting_N=3, reweighting_Taylor=4)
    z = O.zolotarev(3, np.sqrt(0.02), np.sqrt(70.0))
    dU = dev(U)
    Q.seed(4242, 1, 7)
    W = Q.reweighting_factor(dU, 0.25, 10000, 1e-26, prm, None, rp)
    Wr, st = O.reweighting_factor(U, 0.25, z, 3, 4, 10000, 1e-26, 4242, 1, 7)
    assert st == 0
    assert abs(W - Wr) < 1e-10 * abs(Wr)
    assert abs(W - 1.0) > 1e-8
    # list form: product over fermions, one trajectory index each
    Q.seed(4242, 1, 7)
    rp2 = Q.get_rhmc_params(4, np.sqrt(0.02), np.sqrt(70.0))
    z2 = O.zolotarev(4, np.sqrt(0.02), np.sqrt(70.0))
    Wl = Q.reweighting_factor(dU, [0.25, 0.3], 10000, 1e-26, prm, None, [rp, rp2])
    Wr2, _ = O.reweighting_factor(U, 0.3, z2, 1, 5, 10000, 1e-26, 4242, 1, 8)
    assert abs(Wl - Wr * Wr2) < 1e-10 * abs(Wr * Wr2)


def test_autotune_then_parity(Q, O):
    """The autotuner picks valid tiles and leaves results within parity."""
    say("autotune")
    L1, L2 = 64, 48
    rng = np.random.default_rng(8)
    U = random_links(L1, L2, rng, unitary=True)
    si = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    t = Q.autotune(prm, 12)
    assert set(t) == {"q64", "q32", "cg64", "cg32", "cg64_k2", "cg32_k2"}
    assert all(v in (1, 2, 4, 8, 16) for v in t.values())
    assert Q.get_tiles(prm) == t
    so = Q.spinor(prm)
    Q.gamm5Dw(so, dev(U), dev(si), 0.2, prm)
    assert relerr(host(so), O.gamm5Dw(U, si, 0.2)) < 1e-13
    for prec in (0, 1):
        x = Q.spinor(prm)
        Q.CG(x, dev(U), dev(si), 0.2, 5000, 1e-24, Q.gamm5Dw_sqr_msq, prm, None, precision=prec)
        ref, _, st = O.cg(U, si, 0.2, 5000, 1e-24)
        assert st == 0
        assert relerr(host(x), ref) < 1e-9
    Q.set_tiles(prm, 2, 0, 4, 0)
    assert Q.get_tiles(prm)["q64"] == 2 and Q.get_tiles(prm)["cg64"] == 4


@pytest.mark.parametrize("L1,L2,nf", [(16, 16, 4), (48, 32, 3)])
def test_rhmc_fused_pseudofermions_bit_exact(Q, O, L1, L2, nf):
    """An RHMC trajectory with every pseudofermion's solves together (mode 1:
    the force's and the acceptance's multi-shift solves over all
    pseudofermions, the heatbath's shifted CG pole by pole over the
    pseudofermions that have that pole) against one pseudofermion at a time
    (mode 2): identical bits (energies, accept/reject, links), with different
    masses and pole counts per pseudofermion."""
    import torch
    say("RHMC fused pseudofermions", L1, L2, nf)
    prm = Q.LattParm((L1, L2), 4.0)
    U0 = Q.spinor(prm)
    Q.hot_start(U0, prm, 21)
    am0 = [0.1, 0.2, 0.15, 0.1][:nf]
    rps = [Q.get_rhmc_params(n, np.sqrt(0.02), np.sqrt(70.0)) for n in (16, 8, 12, 16)[:nf]]
    out = {}
    for mode in (1, 2):
        U = U0.clone()
        Q.set_multishift(prm, mode)
        Q.seed(33)
        infos = [Q.HMC(U, am0, 0.1, 3, [], 100000, 1e-22, prm, None, rps, integrator=Q.OMF4())
                 for _ in range(2)]
        out[mode] = (infos, U)
    Q.set_multishift(prm, 1)
    for a, b in zip(out[1][0], out[2][0]):
        assert a.hini == b.hini and a.hfin == b.hfin and a.accepted == b.accepted
    assert torch.equal(out[1][1], out[2][1])


@pytest.mark.parametrize("unitary", [True, False])
def test_mixed_cg_packed_links_and_fallback(Q, O, unitary):
    """The mixed solver's site-pair kernels read U(1) links packed into one
    float each; links that are not of unit modulus (here: random complex)
    make the first look at the state restart the solve with FP32 complex
    links.  Either way the FP64 true residual meets the tolerance and the
    solution matches the FP64 solve; the batched solver agrees bit for bit."""
    import torch
    say("mixed CG packed links, unitary", unitary)
    L1, L2 = 48, 32
    rng = np.random.default_rng(9 + unitary)
    U = random_links(L1, L2, rng, unitary=unitary)
    b = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    dU, db = dev(U), dev(b)
    x1, x0 = Q.spinor(prm), Q.spinor(prm)
    it1 = Q.CG(x1, dU, db, 0.3, 10000, 1e-22, Q.gamm5Dw_sqr_msq, prm, None, precision=1)
    it0 = Q.CG(x0, dU, db, 0.3, 10000, 1e-22, Q.gamm5Dw_sqr_msq, prm, None, precision=0)
    assert _true_res(O, U, host(x1), b, 0.3) < 1e-22
    assert relerr(host(x1), host(x0)) < 1e-9
    say("iterations mixed", it1, "fp64", it0)
    assert it1 < 1.5 * it0 + 20
    # the same solve through a right-hand-side set of one, against the
    # single-lattice solve, both with host-decided updates (a set follows its
    # context's mode)
    Q.set_reliable_mode(prm, False)
    x2 = Q.spinor(prm)
    it2 = Q.CG(x2, dU, db, 0.3, 10000, 1e-22, Q.gamm5Dw_sqr_msq, prm, None, precision=1)
    assert relerr(host(x2), host(x0)) < 1e-9
    xs = torch.zeros((1, 2, L2, L1), dtype=torch.complex128, device="cuda")
    its = Q.CG_mrhs(xs, dU, db.reshape(1, 2, L2, L1).contiguous(), 0.3, 10000, 1e-22, prm, None,
                    precision=1)
    Q.set_reliable_mode(prm, True)
    assert its[0] == it2 and torch.equal(xs[0], x2)


@pytest.mark.parametrize("kind", ["pure", "nf2"])
def test_qzero_veto(Q, O, kind):
    """HMC!'s Qzero option (src/QED2dHMC.jl:31-39,86-93): a trajectory ending
    with |Q| > 0.1 is rejected whatever its dH and the links are restored; from
    a cold start (Q = 0 throughout) the option changes nothing.  24x24, beta 4,
    eps 0.02, ns 10 (OMF4 pure gauge, leapfrog Nf=2 at m0 0.3), against the
    oracle on the same streams."""
    L, seed = 24, 77
    p = O.latt(L, L, 4.0)
    prm = Q.LattParm((L, L), 4.0)
    res = {}
    for start in ("hot", "cold"):
        U0 = O.hot_start(L, L, seed) if start == "hot" else np.ones((2, L, L), complex)
        for qz in (False, True):
            say("qzero", kind, start, qz)
            U, dU = U0.copy(), dev(U0)
            Q.seed(seed)
            if kind == "pure":
                info = Q.HMC(dU, 0.02, 10, [], prm, None, qzero=qz)
                a = O.hmc_pure(U, 0.02, 10, p, seed, 0, qzero=int(qz))[2]
            else:
                info = Q.HMC(dU, 0.3, 0.02, 10, [], 10000, 1e-20, prm, None, qzero=qz)
                _, _, a, _, rc = O.hmc_nf2(U, 0.3, 0.02, 10, 10000, 1e-20, p, seed, 0,
                                           qzero=int(qz))
                assert rc == 0
            assert info.accepted == a, (start, qz)
            assert relerr(host(dU), U) < 1e-9
            res[(start, qz)] = (a, O.qtop(U))
    q0 = O.qtop(O.hot_start(L, L, seed))
    assert abs(q0) > 0.1
    # the hot start is accepted without the option and vetoed with it
    assert res[("hot", False)][0] == 1 and abs(res[("hot", False)][1]) > 0.1
    assert res[("hot", True)][0] == 0 and abs(res[("hot", True)][1] - q0) < 1e-9
    # Q = 0 from a cold start: the same decision either way
    assert res[("cold", True)][0] == res[("cold", False)][0]
    assert abs(res[("cold", True)][1]) < 1e-9
