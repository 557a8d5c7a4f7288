"""Workload-level parity: the benchmark configurations of the path, run through
the C ABI at their real shapes and compared with the FP64 oracle where the
oracle finishes in seconds, and through size-independent properties where it
does not (true residuals, reversibility, integer topological charge, mixed vs
FP64 agreement).

Tolerances, stated per test:
* Nf=2 24x24: identical accept/reject and |dH_gpu - dH_oracle| < 1e-9 over 20
  trajectories, FP64 and mixed, at CG tol 1e-24 and at the reference's 1e-12
  (measured on MI355X: at most 9e-13 in FP64 and 3.2e-11 mixed).
* 16-RHS CG at 1024x1024: FP64 and mixed both reach a true FP64 residual
  |b - Q^2 x|^2 < 1e-12 |b|^2; the two solutions agree to 1e-4 relative; the
  16 solves in one set equal the one-at-a-time solves bit for bit.
* master field: per-site plaquette and charge densities within 1e-13 of the
  oracle, Q an integer to 1e-6 and equal to the oracle's after rounding.
"""
import numpy as np
import pytest

from gpu_util import dev, host, relerr, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc,prec,tol", [(0, 0, 1e-24), (1, 0, 1e-24), (0, 1, 1e-24),
                                         (1, 1, 1e-24), (0, 0, 1e-12), (1, 0, 1e-12)])
def test_nf2_24x24_twenty_trajectories(Q, O, bc, prec, tol):
    """Nf=2 HMC, 24x24, beta 4, m0 0.2, hot start seed 1234, leapfrog ns 10
    eps 0.1, 20 trajectories; periodic and antiperiodic fermion time BC."""
    say("nf2 24x24 20 traj bc", bc, "prec", prec, "tol", tol)
    L1 = L2 = 24
    prm = Q.LattParm((L1, L2), 4.0, bc)
    U = O.hot_start(L1, L2, 1234)
    dU = dev(U)
    p = O.latt(L1, L2, 4.0, bc)
    Q.seed(1234)
    acc = []
    dmax = 0.0
    n_acc = 0
    for t in range(20):
        info = Q.HMC(dU, 0.2, 0.1, 10, acc, 100000, tol, prm, None, precision=prec)
        hi, hf, a, it, rc = O.hmc_nf2(U, 0.2, 0.1, 10, 100000, tol, p, 1234, t)
        assert rc == 0
        assert info.accepted == a, t
        dmax = max(dmax, abs((info.hfin - info.hini) - (hf - hi)))
        n_acc += a
    say("max |ddH|", dmax, "accepted", n_acc)
    assert dmax < 1e-9
    assert 0 < n_acc < 20
    assert relerr(host(dU), U) < 1e-9


def _thermalised_1024(Q, beta, n):
    prm = Q.LattParm((1024, 1024), beta)
    U = Q.cold_links(prm)
    Q.seed(1234, chain=7)
    for _ in range(n):
        Q.HMC(U, 0.1, 10, [], prm, None)
    return prm, U


def test_nf2_1024_dh_vs_oracle(Q, O):
    """The headline trajectory: Nf=2 at 1024x1024, beta 4, m0 0.2, leapfrog
    ns 10 eps 0.1, CG tol 1e-24, links after 200 pure-gauge trajectories.
    Mixed precision (FP32 force solves, FP64 reliable updates) and FP64 on the
    GPU against the FP64 oracle on the same links and random streams, hfin
    from the last MD solve as the reference: the same accept/reject and
    |dH - dH_oracle| < 1e-9 (both dH from energies kept below the ulp of H,
    which is ~5e-10 at H ~ 3e6)."""
    import torch
    say("nf2 1024^2 dH vs oracle")
    prm, U = _thermalised_1024(Q, 4.0, 200)
    Uh = host(U).copy()
    p = O.latt(1024, 1024, 4.0)
    hi, hf, a, it, rc = O.hmc_nf2(Uh, 0.2, 0.1, 10, 100000, 1e-24, p, 1234, 0)
    assert rc == 0
    dh_o = O.last_dh()
    say("oracle dH", dh_o, "iterations", it)
    U0 = U.clone()
    for prec in (1, 0):
        U.copy_(U0)
        Q.seed(1234)
        info = Q.HMC(U, 0.2, 0.1, 10, [], 100000, 1e-24, prm, None, precision=prec)
        torch.cuda.synchronize()
        say("prec", prec, "dH", info.dH, "diff", info.dH - dh_o, "it", info.cg_iters)
        assert info.accepted == a
        assert abs(info.dH - dh_o) < 1e-9


def test_cg_16rhs_1024_fp64_vs_mixed(Q, O):
    """Links from 200 pure-gauge OMF4 trajectories at beta 4 from a cold start,
    16 complex Gaussian right-hand sides, m0 0.1, CG to |r|^2 < 1e-12 |b|^2
    in FP64 and in FP32 with FP64 reliable updates; the 16 solves at once
    (CG_mrhs) equal the one-at-a-time solves bit for bit."""
    import torch
    say("16-RHS CG 1024^2")
    prm, U = _thermalised_1024(Q, 4.0, 200)
    nrhs = 16
    its = {0: [], 1: []}
    si = torch.zeros((nrhs, 2, 1024, 1024), dtype=torch.complex128, device="cuda")
    for j in range(nrhs):
        Q.fill_pseudofermion(si[j], prm, 1234, 3, 0, j)
    xs = {0: torch.zeros_like(si), 1: torch.zeros_like(si)}
    for j in range(nrhs):
        b = si[j]
        for prec in (0, 1):
            x = xs[prec][j]
            it = Q.CG(x, U, b, 0.1, 100000, 1e-12, Q.gamm5Dw_sqr_msq, prm, None, precision=prec)
            its[prec].append(it)
            ax, tmp = Q.spinor(prm), Q.spinor(prm)
            Q.gamm5Dw_sqr_msq(ax, tmp, U, x, 0.1, prm)
            r2 = float(torch.sum(torch.abs(b - ax) ** 2))
            b2 = float(torch.sum(torch.abs(b) ** 2))
            assert r2 < 1e-12 * b2, (j, prec, r2 / b2)
        d = float(torch.sum(torch.abs(xs[0][j] - xs[1][j]) ** 2) /
                  torch.sum(torch.abs(xs[0][j]) ** 2))
        assert d < 1e-8, (j, d)
    say("iterations fp64", its[0], "mixed", its[1])
    assert max(i1 - 1.3 * i0 for i0, i1 in zip(its[0], its[1])) < 20
    # the set against one-at-a-time solves, both with reliable updates decided
    # by the host (a set follows its context's mode)
    Q.set_reliable_mode(prm, False)
    ref = {0: (its[0], xs[0]), 1: ([], torch.zeros_like(si))}
    for j in range(nrhs):
        ref[1][0].append(Q.CG(ref[1][1][j], U, si[j], 0.1, 100000, 1e-12, Q.gamm5Dw_sqr_msq,
                              prm, None, precision=1))
    for prec in (0, 1):
        say("CG_mrhs 16 x 1024^2 prec", prec)
        so = torch.zeros_like(si)
        itm = Q.CG_mrhs(so, U, si, 0.1, 100000, 1e-12, prm, None, precision=prec)
        assert itm == ref[prec][0], prec
        assert torch.equal(so, ref[prec][1]), prec
    Q.set_reliable_mode(prm, True)


@pytest.mark.parametrize("L", [1024])
def test_pure_gauge_master_field_trajectories(Q, O, L):
    """Pure-gauge OMF4 HMC at beta 11.25, ns 20, eps 0.025 from a hot start:
    two trajectories against the oracle (energies, accept/reject, links)."""
    say("pure gauge master-field params", L)
    prm = Q.LattParm((L, L), 11.25)
    U = O.hot_start(L, L, 99)
    dU = dev(U)
    p = O.latt(L, L, 11.25)
    Q.seed(99)
    for t in range(2):
        info = Q.HMC(dU, 0.025, 20, [], prm, None)
        hi, hf, a = O.hmc_pure(U, 0.025, 20, p, 99, t)
        assert info.accepted == a
        assert abs(info.hini - hi) < 1e-6 * abs(hi)
        assert abs((info.hfin - info.hini) - (hf - hi)) < 1e-6
    assert relerr(host(dU), U) < 1e-9


def test_master_field_8192_densities(Q, O):
    """8192x8192 pure gauge at beta 11.25 (ns 20, eps 0.025): after two GPU
    trajectories the per-site plaquette and topological-charge densities match
    the oracle on the same links, and Q is an integer equal to the oracle's."""
    say("master field 8192^2")
    L = 8192
    prm = Q.LattParm((L, L), 11.25)
    U = Q.spinor(prm)
    Q.hot_start(U, prm, 4321)
    Q.seed(4321)
    for _ in range(2):
        Q.HMC(U, 0.025, 20, [], prm, None)
    plx, qx = Q.realfield(prm)[0], Q.realfield(prm)[0]
    Q.plaquette(plx, U, prm)
    Q.qtop(qx, U, prm)
    qg = Q.Qtop(U, prm)
    Uh = host(U)
    pr, qr = O.plaquette_field(Uh), O.qtop_field(Uh)
    assert np.max(np.abs(host(plx) - pr)) < 1e-13
    assert np.max(np.abs(host(qx) - qr)) < 1e-13
    qo = O.qtop(Uh)
    say("Q gpu", qg, "oracle", qo)
    assert abs(qg - round(qg)) < 1e-6
    assert round(qg) == round(qo)
    assert abs(Q.Plaquette(U, prm) - O.plaquette(Uh)) < 1e-12


def test_nf2_4096_mixed_vs_fp64(Q, O):
    """Nf=2 at 4096x4096, beta 11.25, m0 0.05 (links after 3 pure-gauge
    trajectories from a cold start): one leapfrog trajectory (ns 2, eps 0.05)
    with FP32 force solves and reliable updates against the FP64 form at the
    same CG tolerance 1e-20: same accept/reject, and |dH_mixed - dH_fp64| below
    a tenth of what the tolerance itself moves dH (FP64 at 1e-20 against FP64
    at 1e-22, about 1e-6 here) and below 1e-7."""
    import torch
    say("nf2 4096^2 mixed vs fp64")
    L = 4096
    prm = Q.LattParm((L, L), 11.25)
    U = Q.cold_links(prm)
    Q.seed(55, chain=2)
    for _ in range(3):
        Q.HMC(U, 0.025, 20, [], prm, None)
    U0 = U.clone()
    res = {}
    for key, prec, tol in ((1, 1, 1e-20), (0, 0, 1e-20), ("tight", 0, 1e-22)):
        U.copy_(U0)
        Q.seed(55)
        res[key] = Q.HMC(U, 0.05, 0.05, 2, [], 100000, tol, prm, None, precision=prec,
                         acc_tol=tol)
        torch.cuda.synchronize()
    moved = abs(res[0].dH - res["tight"].dH)
    say("dH mixed", res[1].dH, "fp64", res[0].dH, "it", res[1].cg_iters, res[0].cg_iters,
        "fp64 1e-20 vs 1e-22", moved)
    assert res[1].accepted == res[0].accepted
    assert abs(res[1].dH - res[0].dH) < min(0.1 * moved, 1e-7)


def test_rhmc_128_sixteen_poles_four_pseudofermions(Q, O):
    """RHMC at 128x128, beta 4, m0 0.1: 16 Zolotarev poles on spectral bounds
    from the power method (x0.99 / x1.01), 4 pseudofermions, OMF4 ns 12 eps
    0.083.  Reversibility (test/rhmc_hmc_reversibility.jl) and a bounded energy
    violation; the multi-shift force agrees with the sequential one."""
    import torch
    say("RHMC 128^2 16 poles 4 pf")
    L = 128
    prm = Q.LattParm((L, L), 4.0)
    U = Q.cold_links(prm)
    Q.seed(77, chain=3)
    for _ in range(20):
        Q.HMC(U, 0.1, 10, [], prm, None)
    b0 = Q.spinor(prm)
    Q.fill_pseudofermion(b0, prm, 77, 3, 0, 0)
    lmin, lmax = Q.power_method(U, 0.1, prm, iter=5000, b0=b0)
    rp = Q.get_rhmc_params(16, np.sqrt(lmin * 0.99), np.sqrt(lmax * 1.01))
    say("spectrum", lmin, lmax, "delta", rp.delta)
    U0 = U.clone()
    S0 = Q.Action(U, prm)
    Q.seed(77)
    info = Q.HMC(U, [0.1] * 4, 0.083, 12, [], 100000, 1e-22, prm, None, [rp] * 4,
                 integrator=Q.OMF4(), reversibility_test=True)
    assert info.accepted == -1
    assert abs(Q.Action(U, prm) - S0) < 1e-9
    out = {}
    for ms in (1, 0):
        U.copy_(U0)
        Q.set_multishift(prm, ms)
        Q.seed(77)
        out[ms] = Q.HMC(U, [0.1] * 4, 0.083, 12, [], 100000, 1e-22, prm, None, [rp] * 4,
                        integrator=Q.OMF4())
    Q.set_multishift(prm, 1)
    # the force's solves in mixed precision (FP32 multi-shift, then every pole
    # by the batched mixed-precision CG) to the same FP64 tolerance
    U.copy_(U0)
    Q.seed(77)
    out["mixed"] = Q.HMC(U, [0.1] * 4, 0.083, 12, [], 100000, 1e-22, prm, None, [rp] * 4,
                         integrator=Q.OMF4(), precision=Q.PREC_MIXED)
    say("dH multishift", out[1].dH, "sequential", out[0].dH, "mixed", out["mixed"].dH)
    assert abs(out[1].dH) < 1.0
    assert out[1].accepted == out[0].accepted == out["mixed"].accepted
    assert abs(out[1].dH - out[0].dH) < 1e-8
    assert abs(out["mixed"].dH - out[1].dH) < 1e-8


@pytest.mark.parametrize("precision", [0, 1])
def test_rhmc_4pf_16poles_omf4_vs_oracle(Q, O, precision):
    """The RHMC configuration's structure (16 poles, 4 pseudofermions, OMF4)
    against the oracle on a 16x16 lattice, one trajectory of 3 steps; precision
    1: the force's solves in mixed precision, at the FP64 tolerances."""
    say("RHMC 16 poles 4 pf OMF4 vs oracle, precision", precision)
    L = 16
    prm = Q.LattParm((L, L), 4.0)
    U = O.hot_start(L, L, 5)
    dU = dev(U)
    p = O.latt(L, L, 4.0)
    rp = Q.get_rhmc_params(16, np.sqrt(0.02), np.sqrt(70.0))
    z = O.zolotarev(16, np.sqrt(0.02), np.sqrt(70.0))
    Q.seed(5)
    info = Q.HMC(dU, [0.1] * 4, 0.083, 3, [], 100000, 1e-24, prm, None, [rp] * 4,
                 integrator=Q.OMF4(), precision=precision)
    hi, hf, a, rc = O.hmc_rhmc(U, [0.1] * 4, [z] * 4, 0.083, 3, 100000, 1e-24, p, 5, 0,
                               integrator=1)
    assert rc == 0
    assert abs(info.hini - hi) < 1e-9 * abs(hi)
    assert abs(info.dH - (hf - hi)) < 1e-8
    assert info.accepted == a


def test_rhmc_32_16poles_4pf_vs_oracle(Q, O):
    """RHMC with 16 poles and 4 pseudofermions against the oracle on a 32x32
    lattice (one OMF4 trajectory of 3 steps, CG tol 1e-24), the force's solves
    in FP64 and in mixed precision, both against one oracle trajectory."""
    say("RHMC 32^2 16 poles 4 pf OMF4 vs oracle")
    L = 32
    U = O.hot_start(L, L, 5)
    p = O.latt(L, L, 4.0)
    z = O.zolotarev(16, np.sqrt(0.02), np.sqrt(70.0))
    hi, hf, a, rc = O.hmc_rhmc(U.copy(), [0.1] * 4, [z] * 4, 0.083, 3, 100000, 1e-24, p, 5,
                               0, integrator=1)  # (the oracle evolves its links in place)
    assert rc == 0
    rp = Q.get_rhmc_params(16, np.sqrt(0.02), np.sqrt(70.0))
    prm = Q.LattParm((L, L), 4.0)
    for precision in (0, 1):
        dU = dev(U)
        Q.seed(5)
        info = Q.HMC(dU, [0.1] * 4, 0.083, 3, [], 100000, 1e-24, prm, None, [rp] * 4,
                     integrator=Q.OMF4(), precision=precision)
        say("precision", precision, "dH", info.dH, "oracle", hf - hi)
        assert abs(info.hini - hi) < 1e-9 * abs(hi)
        assert abs(info.dH - (hf - hi)) < 1e-8
        assert info.accepted == a
