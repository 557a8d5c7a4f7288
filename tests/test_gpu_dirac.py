"""GPU parity of the Wilson operator family, forces, link update, observables
and random fields: libqed2d_hip (through its C ABI, via the qed2d mirror)
against the FP64 CPU oracle on identical inputs.

Tolerances: FP64 kernels 1e-13 relative (max norm) as required of Dw and the
forces; FP32 kernels 2e-6 relative against the FP64 oracle; random fields
bit-exact."""
import numpy as np
import pytest

from dense import random_links, random_spinor
from gpu_util import dev, host, relerr, say

pytestmark = pytest.mark.gpu

SIZES = [(24, 24), (10, 10), (64, 48), (130, 66), (256, 256)]


@pytest.mark.parametrize("L1,L2", SIZES)
@pytest.mark.parametrize("bc", [0, 1])
def test_gamm5Dw_f64(Q, O, L1, L2, bc):
    say("gamm5Dw f64", L1, L2, bc)
    rng = np.random.default_rng(L1 * 1000 + L2 + bc)
    U = random_links(L1, L2, rng)
    si = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0, bc)
    ref = O.gamm5Dw(U, si, 0.2, bc)
    dU, dsi = dev(U), dev(si)
    for rows in (1, 2, 4, 8, 16):
        assert Q.lib.qed2d_ctx_set_tile(prm.ctx(), rows) == 0
        so = Q.spinor(prm)
        Q.gamm5Dw(so, dU, dsi, 0.2, prm)
        assert relerr(host(so), ref) < 1e-13, rows


def test_gamm5Dw_large(Q, O):
    L1, L2 = 1024, 1024
    say("gamm5Dw f64 1024^2")
    rng = np.random.default_rng(7)
    U = random_links(L1, L2, rng, unitary=True)
    si = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    so = Q.spinor(prm)
    Q.gamm5Dw(so, dev(U), dev(si), 0.1, prm)
    assert relerr(host(so), O.gamm5Dw(U, si, 0.1)) < 1e-13


@pytest.mark.parametrize("L1,L2,bc", [(4096, 4096, 1), (4100, 4096, 0)])
def test_gamm5Dw_launch_order(Q, O, L1, L2, bc):
    """From 4096^2 sites on the plain applies run their blocks in launch order
    instead of XCD bands: FP64 and FP32 at two tile heights, against the
    oracle, on a square lattice and on one whose width is not a multiple of
    the block width."""
    say("gamm5Dw launch order", L1, L2, bc)
    rng = np.random.default_rng(L1 + bc)
    U = random_links(L1, L2, rng, unitary=True)
    si = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0, bc)
    ref = O.gamm5Dw(U, si, 0.1, bc)
    dU, dsi = dev(U), dev(si)
    dU32, dsi32 = dev(U.astype(np.complex64)), dev(si.astype(np.complex64))
    for rows in (2, 8):
        Q.set_tiles(prm, q64=rows, q32=rows)
        so = Q.spinor(prm)
        Q.gamm5Dw(so, dU, dsi, 0.1, prm)
        assert relerr(host(so), ref) < 1e-13, rows
        so32 = dev(np.zeros_like(si, dtype=np.complex64))
        Q.gamm5Dw_f32(so32, dU32, dsi32, 0.1, prm)
        assert relerr(host(so32), ref) < 2e-6, rows
    prm.close()


@pytest.mark.parametrize("L1,L2,bc", [(24, 24, 0), (128, 64, 1), (4, 6, 0), (25, 16, 1),
                                       (2, 8, 0), (66, 33, 0)])
def test_gamm5Dw_f32(Q, O, L1, L2, bc):
    """Even L1 >= 4 runs the two-sites-per-thread kernel, odd L1 (and L1 = 2)
    the one-site kernel; every tile height."""
    import torch
    say("gamm5Dw f32", L1, L2, bc)
    rng = np.random.default_rng(3 + L1)
    U = random_links(L1, L2, rng, unitary=True)
    si = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0, bc)
    ref = O.gamm5Dw(U, si, 0.2, bc)
    dU, dsi = dev(U.astype(np.complex64)), dev(si.astype(np.complex64))
    for rows in (1, 2, 4, 8, 16):
        assert Q.lib.qed2d_ctx_set_tile_f32(prm.ctx(), rows) == 0
        so = torch.zeros((2, L2, L1), dtype=torch.complex64, device="cuda")
        Q.gamm5Dw_f32(so, dU, dsi, 0.2, prm)
        assert relerr(host(so), ref) < 2e-6, rows


@pytest.mark.parametrize("k,L1,L2,bc", [(1, 64, 32, 0), (3, 64, 32, 1), (16, 64, 32, 0),
                                        (5, 33, 20, 1), (7, 130, 18, 0)])
def test_gamm5Dw_mrhs(Q, O, k, L1, L2, bc):
    """Every tile height; FP64 and FP32 (pair kernel for even L1, one-site
    kernel for odd L1)."""
    say("gamm5Dw mrhs", k, L1, L2, bc)
    rng = np.random.default_rng(k + L1)
    U = random_links(L1, L2, rng, unitary=True)
    si = np.stack([random_spinor(L1, L2, rng) for _ in range(k)])
    prm = Q.LattParm((L1, L2), 4.0, bc)
    refs = [O.gamm5Dw(U, si[j], 0.15, bc) for j in range(k)]
    dU, dsi = dev(U), dev(si)
    dU32, dsi32 = dev(U.astype(np.complex64)), dev(si.astype(np.complex64))
    for rows in (1, 2, 4, 8, 16):
        Q.set_tiles(prm, q64=rows, q32=rows)
        so = dev(np.zeros_like(si))
        Q.gamm5Dw_mrhs(so, dU, dsi, 0.15, prm)
        out = host(so)
        so32 = dev(np.zeros_like(si, dtype=np.complex64))
        Q.gamm5Dw_mrhs(so32, dU32, dsi32, 0.15, prm)
        out32 = host(so32)
        for j in range(k):
            assert relerr(out[j], refs[j]) < 1e-13, (rows, j)
            assert relerr(out32[j], refs[j]) < 2e-6, (rows, j)


def test_operator_compositions(Q, O):
    say("Dw / gamm5 / sqr compositions")
    L1, L2 = 32, 24
    rng = np.random.default_rng(5)
    U = random_links(L1, L2, rng)
    si = random_spinor(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    dU, dsi = dev(U), dev(si)
    so, tmp = Q.spinor(prm), Q.spinor(prm)
    Q.Dw(so, dU, dsi, 0.3, prm)
    assert relerr(host(so), O.Dw(U, si, 0.3)) < 1e-13
    Q.gamm5(so, dsi, 0.3, prm)
    assert relerr(host(so), O.gamm5(si)) == 0.0
    Q.gamm5Dw_sqr(so, dU, dsi, 0.3, prm)
    assert relerr(host(so), O.gamm5Dw_sqr(U, si, 0.3)) < 1e-13
    Q.gamm5Dw_sqr_msq(so, tmp, dU, dsi, 0.3, prm)
    assert relerr(host(so), O.gamm5Dw_sqr(U, si, 0.3)) < 1e-13
    assert relerr(host(tmp), O.gamm5Dw(U, si, 0.3)) < 1e-13
    Q.gamm5Dw_sqr_musq(so, tmp, dU, dsi, 0.3, 0.8, prm)
    assert relerr(host(so), O.gamm5Dw_sqr_musq(U, si, 0.3, 0.8)) < 1e-13
    Q.gamm5Dw_sqr_sqr(so, dU, dsi, 0.3, prm)
    assert relerr(host(so), O.gamm5Dw_sqr_sqr(U, si, 0.3)) < 1e-13
    Q.gamm5Dw_sqr_sqr_musq(so, tmp, dU, dsi, 0.3, 0.8, prm)
    assert relerr(host(so), O.gamm5Dw_sqr_sqr_musq(U, si, 0.3, 0.8)) < 1e-13


@pytest.mark.parametrize("L1,L2,bc", [(24, 24, 0), (40, 18, 1), (256, 128, 0)])
def test_forces(Q, O, L1, L2, bc):
    say("forces", L1, L2, bc)
    rng = np.random.default_rng(L1 + L2)
    U = random_links(L1, L2, rng)
    X = random_spinor(L1, L2, rng)
    Y = O.gamm5Dw(U, X, 0.2, bc)
    prm = Q.LattParm((L1, L2), 3.5, bc)
    dU, dX, dY = dev(U), dev(X), dev(Y)
    frc = Q.realfield(prm)
    Q.tr_dQwdU(frc, dU, dX, dY, prm)
    fref = O.tr_dQwdU(U, X, Y, bc)
    assert relerr(host(frc), fref) < 1e-13
    f1, f2 = Q.realfield(prm), Q.realfield(prm)
    Q.krnl_force(f1, f2, dU, prm)
    r1, r2 = O.krnl_force(U, 3.5)
    assert relerr(host(f1), r1) < 1e-13 and relerr(host(f2), r2) < 1e-13
    # fused gauge + fermion force + momentum update
    mom0 = rng.standard_normal((2, L2, L1))
    dm = dev(mom0)
    Q.force_mom_update(dm, dU, dX, dY, 0.05, prm)
    ref = mom0 + 0.05 * (r1 + r2 + fref)
    assert relerr(host(dm) - mom0, ref - mom0) < 1e-12


@pytest.mark.parametrize("L1,L2", [(8, 8), (24, 24), (33, 20), (64, 7), (100, 37),
                                   (256, 128), (1024, 512)])
def test_gauge_force_mom(Q, O, L1, L2):
    """The pure-gauge force + momentum update (the marching kernel, mom +=
    eps (frc1 + frc2)): widths below, at and above a wave, odd widths, row
    counts that the tiles do not divide, against the oracle's krnl_force."""
    say("gauge force + mom", L1, L2)
    rng = np.random.default_rng(7 * L1 + L2)
    U = random_links(L1, L2, rng, unitary=(L1 % 2 == 0))
    prm = Q.LattParm((L1, L2), 2.75)
    r1, r2 = O.krnl_force(U, 2.75)
    mom0 = rng.standard_normal((2, L2, L1))
    dm = dev(mom0)
    Q.force_mom_update(dm, dev(U), None, None, 0.07, prm)
    ref = mom0 + 0.07 * (r1 + r2)
    assert relerr(host(dm) - mom0, ref - mom0) < 1e-13
    assert relerr(host(dm), ref) < 1e-15


def test_updt_fld_and_observables(Q, O):
    say("updt_fld / observables")
    L1, L2 = 96, 80
    rng = np.random.default_rng(1)
    U = random_links(L1, L2, rng, unitary=True)
    mom = rng.standard_normal((2, L2, L1))
    prm = Q.LattParm((L1, L2), 4.0)
    dU, dm = dev(U), dev(mom)
    Q.updt_fld(dU, dm, 0.1, prm)
    U2 = O.updt_fld(U, mom, 0.1)
    assert relerr(host(dU), U2) < 1e-15
    # steps beyond pi/4 (the general sincos branch) mixed with small ones
    dU3 = dev(U)
    Q.updt_fld(dU3, dm, 2.5, prm)
    assert relerr(host(dU3), O.updt_fld(U, mom, 2.5)) < 1e-15
    plx = dev(np.zeros((L2, L1)))
    Q.plaquette(plx, dU, prm)
    assert relerr(host(plx), O.plaquette_field(U2)) < 1e-14
    Q.qtop(plx, dU, prm)
    assert relerr(host(plx), O.qtop_field(U2)) < 1e-14
    assert abs(Q.Action(dU, prm) - O.action(U2, 4.0)) < 1e-9
    assert abs(Q.Plaquette(dU, prm) - O.plaquette(U2)) < 1e-14
    assert abs(Q.Qtop(dU, prm) - O.qtop(U2)) < 1e-11
    assert abs(Q.Hamiltonian(dm, dU, prm) - O.hamiltonian(mom, U2, 4.0)) < 1e-9


def test_random_fields_bit_exact(Q, O):
    say("random fields")
    L1, L2 = 48, 40
    prm = Q.LattParm((L1, L2), 4.0)
    mom = Q.realfield(prm)
    Q.fill_momenta(mom, prm, 1234, 3, 17)
    assert np.array_equal(host(mom), O.momenta(L1, L2, 1234, 3, 17))
    X = Q.spinor(prm)
    Q.fill_pseudofermion(X, prm, 98765432101, 0, 5, 2)
    assert np.array_equal(host(X), O.pseudofermion(L1, L2, 98765432101, 0, 5, 2))
    U = Q.spinor(prm)
    Q.hot_start(U, prm, 1234, 0)
    assert np.array_equal(host(U), O.hot_start(L1, L2, 1234, 0))
    for t in range(5):
        assert Q.metropolis_uniform(1234, 0, t) == O.metropolis_uniform(1234, 0, t)


def test_unfold(Q, O):
    say("unfold")
    L1, L2 = 16, 12
    rng = np.random.default_rng(0)
    U = random_links(L1, L2, rng)
    prm = Q.LattParm((L1, L2), 4.0)
    assert np.array_equal(host(Q.unfold_fld(dev(U), prm)), O.unfold(U, False))
    assert np.array_equal(host(Q.unfold_reflect(dev(U), prm)), O.unfold(U, True))
