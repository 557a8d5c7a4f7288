"""Batched ensembles (qed2d_batch_*): a chain of a batch against the
single-lattice entry points on the same lattice, seeds and tiles.  The batch
runs the same device code per chain with the same tiling, so the comparison
is bit-exact (CG solutions, energies, accept/reject, links), FP64 and mixed
precision (the batch with its chains pausing for reliable updates, the
single lattice with host-decided updates: the same schedule).  In its default
mode a batch's chains request updates and iterate on; that mode is checked
against FP64 solves and, in the ensemble of 64 chains of uneven sizes, against
FP64 trajectories (accept/reject identical, |dH_mixed - dH_fp64| < 1e-9)."""
import numpy as np
import pytest

from gpu_util import dev, host, relerr, say

pytestmark = pytest.mark.gpu

TILES = dict(q64=4, cg64=16, cg32=8)


def _single(Q, L1, L2, beta, bc=0):
    prm = Q.LattParm((L1, L2), beta, bc)
    Q.set_tiles(prm, q64=TILES["q64"], cg64=TILES["cg64"], cg32=TILES["cg32"])
    Q.set_reliable_mode(prm, False)
    return prm


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_cg_bit_exact(Q, O, prec):
    import torch
    say("batch CG prec", prec)
    sizes = [(16, 16), (32, 24), (64, 64), (30, 18)]
    betas = [2.0, 4.0, 6.0, 4.0]
    am0s = [0.1, 0.2, 0.15, 0.3]
    B = Q.Batch(sizes, betas)
    B.set_reliable_mode(False)  # (the single lattice's host-decided schedule)
    B.set_tiles(TILES["q64"], TILES["cg64"], TILES["cg32"])
    B.hot_start([11, 12, 13, 14])
    src = B.arena_spinor()
    for c, (L1, L2) in enumerate(sizes):
        prm = _single(Q, L1, L2, betas[c])
        Q.fill_pseudofermion(B.field(src, c), prm, 500 + c, c, 0, 0)
    x = B.arena_spinor()
    its = B.CG(x, src, am0s, 10000, 1e-20, precision=prec)
    for c, (L1, L2) in enumerate(sizes):
        prm = _single(Q, L1, L2, betas[c])
        U = B.get_links(c)
        U1 = Q.spinor(prm)
        Q.hot_start(U1, prm, 11 + c, chain=c)
        assert torch.equal(U, U1)
        xs = Q.spinor(prm)
        it = Q.CG(xs, U, B.field(src, c).contiguous(), am0s[c], 10000, 1e-20,
                  Q.gamm5Dw_sqr_msq, prm, None, precision=prec)
        assert it == its[c], (c, it, its[c])
        assert torch.equal(xs, B.field(x, c)), c


def test_batch_cg_soft_reliable_updates(Q, O):
    """Mixed-precision batched CG in its default mode (a chain requests a
    reliable update and iterates on until the next chunk start) against FP64
    single-lattice solves: the FP64 stopping rule holds on every chain's true
    residual, the solutions agree to what the tolerance allows, and the
    iteration counts stay close to FP64's."""
    say("batch CG soft reliable updates")
    sizes = [(16, 16), (32, 24), (64, 64), (30, 18), (128, 96)]
    betas = [2.0, 4.0, 6.0, 4.0, 4.0]
    am0s = [0.1, 0.2, 0.15, 0.3, 0.1]
    B = Q.Batch(sizes, betas)
    B.hot_start([51, 52, 53, 54, 55])
    src = B.arena_spinor()
    for c, (L1, L2) in enumerate(sizes):
        prm = Q.LattParm((L1, L2), betas[c])
        Q.fill_pseudofermion(B.field(src, c), prm, 700 + c, c, 0, 0)
    x = B.arena_spinor()
    its = B.CG(x, src, am0s, 10000, 1e-22, precision=1)
    for c, (L1, L2) in enumerate(sizes):
        prm = Q.LattParm((L1, L2), betas[c])
        U = B.get_links(c)
        b = B.field(src, c).contiguous()
        xr = Q.spinor(prm)
        it = Q.CG(xr, U, b, am0s[c], 10000, 1e-22, Q.gamm5Dw_sqr_msq, prm, None)
        xc = B.field(x, c).contiguous()
        res = O.norm2(host(b) - O.gamm5Dw_sqr_musq(host(U), host(xc), am0s[c], 0.0))
        assert res < 1e-22 * O.norm2(host(b)), (c, res)
        assert relerr(host(xc), host(xr)) < 1e-8, c
        assert abs(its[c] - it) <= 0.1 * it + 8, (c, its[c], it)


@pytest.mark.parametrize("prec,integ", [(0, "leapfrog"), (1, "leapfrog"), (1, "omf4")])
def test_batch_hmc_bit_exact(Q, O, prec, integ):
    import torch
    say("batch HMC prec", prec, integ)
    sizes = [(16, 16), (32, 32), (24, 16)]
    betas = [2.0, 4.0, 6.0]
    am0s = [0.1, 0.2, 0.1]
    seeds = [100, 101, 102]
    integrator = Q.OMF4() if integ == "omf4" else Q.Leapfrog()
    ns = 3 if integ == "omf4" else 10
    B = Q.Batch(sizes, betas)
    B.set_reliable_mode(False)  # (the single lattice's host-decided schedule)
    B.set_tiles(TILES["q64"], TILES["cg64"], TILES["cg32"])
    B.hot_start(seeds)
    singles = []
    for c, (L1, L2) in enumerate(sizes):
        prm = _single(Q, L1, L2, betas[c])
        singles.append((prm, B.get_links(c).clone()))
    for t in range(3):
        infos = B.HMC(am0s, 0.1, ns, 100000, 1e-22, seeds, t, integrator=integrator,
                      precision=prec)
        for c, (prm, U) in enumerate(singles):
            Q.seed(seeds[c], chain=c, traj=t)
            ref = Q.HMC(U, am0s[c], 0.1, ns, [], 100000, 1e-22, prm, None,
                        integrator=integrator, precision=prec)
            assert infos[c].accepted == ref.accepted, (t, c)
            assert infos[c].hini == ref.hini and infos[c].hfin == ref.hfin, (t, c)
            assert infos[c].cg_iters == ref.cg_iters, (t, c)
            assert torch.equal(B.get_links(c), U), (t, c)
    p, q, a = B.observables()
    for c, (prm, U) in enumerate(singles):
        assert p[c] == Q.Plaquette(U, prm) and a[c] == Q.Action(U, prm)


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_hmc_graphs_bit_exact(Q, O, prec):
    """CG chunks queued as HIP graph launches (set_graphs) against direct
    launches: the same kernels replayed, so the same bits, over three
    trajectories (every chunk shape's graph replayed many times, with the
    trajectories' other work queued directly between the replays)."""
    import torch
    say("batch HMC graphs prec", prec)
    sizes = [(16, 16), (32, 32), (64, 48)]
    betas = [2.0, 4.0, 6.0]
    am0s = [0.1, 0.2, 0.1]
    seeds = [200, 201, 202]
    Bs = []
    for graphs in (False, True):
        B = Q.Batch(sizes, betas)
        B.set_tiles(TILES["q64"], TILES["cg64"], TILES["cg32"])
        B.hot_start(seeds)
        B.set_graphs(graphs)
        Bs.append(B)
    for t in range(3):
        a = Bs[0].HMC(am0s, 0.1, 10, 100000, 1e-22, seeds, t, precision=prec)
        b = Bs[1].HMC(am0s, 0.1, 10, 100000, 1e-22, seeds, t, precision=prec)
        for c in range(len(sizes)):
            assert (a[c].hini, a[c].hfin, a[c].accepted, a[c].cg_iters) == \
                (b[c].hini, b[c].hfin, b[c].accepted, b[c].cg_iters), (t, c)
            assert torch.equal(Bs[0].get_links(c), Bs[1].get_links(c)), (t, c)
    for B in Bs:
        B.close()


def _ensemble():
    sizes = [(16, 16)] * 16 + [(32, 32)] * 16 + [(64, 64)] * 16 + [(128, 128)] * 8 + \
            [(256, 256)] * 8
    betas = [(2.0, 4.0, 6.0)[c % 3] for c in range(64)]
    am0s = [(0.1, 0.2)[(c // 3) % 2] for c in range(64)]
    seeds = [1000 + c for c in range(64)]
    return sizes, betas, am0s, seeds


def test_batch_ensemble_64_chains_mixed_vs_fp64(Q, O):
    """64 chains of uneven sizes (16:16:16:8:8 of 16^2 .. 256^2), beta in
    {2, 4, 6}, m0 in {0.1, 0.2}, each seeded separately; leapfrog ns 10 eps
    0.1; from hot starts and two mixed-precision trajectories, one trajectory
    in mixed precision against the same trajectory in FP64 (FP64 acceptance
    check, CG tol 1e-24; at 1e-22 the largest difference was 1.05e-9)."""
    import torch
    say("64-chain ensemble")
    sizes, betas, am0s, seeds = _ensemble()
    B = Q.Batch(sizes, betas)
    B.hot_start(seeds)
    for t in range(2):
        B.HMC(am0s, 0.1, 10, 100000, 1e-20, seeds, t, precision=1)
    U0 = [B.get_links(c).clone() for c in range(64)]
    res = {}
    for prec in (1, 0):
        for c in range(64):
            B.set_links(c, U0[c])
        res[prec] = B.HMC(am0s, 0.1, 10, 100000, 1e-24, seeds, 2, precision=prec,
                          acc_tol=1e-24)
    dd = [abs(a.dH - b.dH) for a, b in zip(res[1], res[0])]
    say("max |dH_mixed - dH_fp64|", max(dd), "accepted", sum(i.accepted for i in res[1]))
    assert [i.accepted for i in res[1]] == [i.accepted for i in res[0]]
    assert max(dd) < 1e-9


@pytest.mark.parametrize("L1,L2,prec", [(64, 48, 0), (64, 48, 1), (37, 20, 0), (37, 20, 1)])
def test_cg_mrhs_bit_exact(Q, O, L1, L2, prec):
    """CG_mrhs (k right-hand sides sharing the links, one launch per kernel)
    against CG on each right-hand side alone, same context and tiles:
    identical iteration counts and solutions; a second call with fewer
    right-hand sides rebuilds the set.  The odd L1 = 37 mixed case runs the
    right-hand sides one at a time (the set's FP32 kernels march site pairs)."""
    import torch
    say("cg_mrhs", L1, L2, "prec", prec)
    prm = Q.LattParm((L1, L2), 4.0)
    Q.set_reliable_mode(prm, False)
    U = Q.spinor(prm)
    Q.hot_start(U, prm, 21)
    for k in (5, 3):
        si = torch.zeros((k, 2, L2, L1), dtype=torch.complex128, device="cuda")
        for j in range(k):
            Q.fill_pseudofermion(si[j], prm, 31, 0, k, j)
        so = torch.zeros_like(si)
        its = Q.CG_mrhs(so, U, si, 0.15, 10000, 1e-20, prm, None, precision=prec)
        for j in range(k):
            x = Q.spinor(prm)
            it = Q.CG(x, U, si[j].contiguous(), 0.15, 10000, 1e-20, Q.gamm5Dw_sqr_msq, prm,
                      None, precision=prec)
            assert it == its[j], (k, j, it, its[j])
            assert torch.equal(x, so[j]), (k, j)
    # the tiles in effect are the context's (the second CG kernel's included):
    # change them and compare again
    Q.set_tiles(prm, q64=4, cg64=8, cg32=4, cg64_k2=2, cg32_k2=8)
    si = torch.zeros((2, 2, L2, L1), dtype=torch.complex128, device="cuda")
    for j in range(2):
        Q.fill_pseudofermion(si[j], prm, 32, 0, 0, j)
    so = torch.zeros_like(si)
    its = Q.CG_mrhs(so, U, si, 0.15, 10000, 1e-20, prm, None, precision=prec)
    for j in range(2):
        x = Q.spinor(prm)
        assert Q.CG(x, U, si[j].contiguous(), 0.15, 10000, 1e-20, Q.gamm5Dw_sqr_msq, prm, None,
                    precision=prec) == its[j]
        assert torch.equal(x, so[j]), j


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_split_tiles_and_autotune_bit_exact(Q, O, prec):
    """A batch whose CG kernels K1 and K2 run different tile heights (set, then
    chosen by the batch's autotuner) reproduces single-lattice CG with the same
    heights bit for bit; the default heights leave at least 1024 blocks."""
    import torch
    say("batch split tiles / autotune prec", prec)
    sizes = [(16, 16), (32, 24), (64, 64), (30, 18)]
    betas = [2.0, 4.0, 6.0, 4.0]
    am0s = [0.1, 0.2, 0.15, 0.3]
    B = Q.Batch(sizes, betas)
    B.set_reliable_mode(False)  # (the single lattice's host-decided schedule)
    d = B.get_tiles()
    assert all(v in (1, 2, 4, 8, 16) for v in d.values())
    B.hot_start([41, 42, 43, 44])
    src = B.arena_spinor()
    for c, (L1, L2) in enumerate(sizes):
        prm = Q.LattParm((L1, L2), betas[c])
        Q.fill_pseudofermion(B.field(src, c), prm, 600 + c, c, 0, 0)
    for mode in ("split", "autotune"):
        if mode == "split":
            B.set_tiles(2, 4, 2, cg64_k2=16, cg32_k2=8)
        else:
            B.autotune(8)
        t = B.get_tiles()
        say("batch tiles", mode, t)
        x = B.arena_spinor()
        its = B.CG(x, src, am0s, 10000, 1e-20, precision=prec)
        for c, (L1, L2) in enumerate(sizes):
            prm = Q.LattParm((L1, L2), betas[c])
            Q.set_reliable_mode(prm, False)
            Q.set_tiles(prm, q64=t["q64"], cg64=t["cg64"], cg32=t["cg32"],
                        cg64_k2=t["cg64_k2"], cg32_k2=t["cg32_k2"])
            xs = Q.spinor(prm)
            it = Q.CG(xs, B.get_links(c), B.field(src, c).contiguous(), am0s[c], 10000, 1e-20,
                      Q.gamm5Dw_sqr_msq, prm, None, precision=prec)
            assert it == its[c], (mode, c, it, its[c])
            assert torch.equal(xs, B.field(x, c)), (mode, c)
