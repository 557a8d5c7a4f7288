"""qed2d_hamiltonian_noalloc (no context, no allocation; scratch in a static
device array) against the context entry points it mirrors: qed2d_hamiltonian,
qed2d_action, qed2d_plaquette and qed2d_qtop, bit for bit, on two calls in a
row (the second on a smaller lattice, so stale partials of the first call
would show)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def _call(Q, torch, L1, L2, beta, mom, U):
    out = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    rc = Q.lib.qed2d_hamiltonian_noalloc(
        L1, L2, C.c_double(beta), C.c_void_p(mom.data_ptr() if mom is not None else 0),
        C.c_void_p(U.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(st))
    assert rc == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _reference(Q, L1, L2, beta, mom, U):
    prm = Q.LattParm((L1, L2), beta)
    try:
        return np.array([Q.Hamiltonian(mom, U, prm), Q.Action(U, prm), Q.Plaquette(U, prm),
                         Q.Qtop(U, prm)])
    finally:
        prm.close()


def _fields(Q, torch, L1, L2, seed):
    prm = Q.LattParm((L1, L2), 1.0)
    U = Q.spinor(prm)
    Q.hot_start(U, prm, seed)
    mom = torch.zeros((2, L2, L1), dtype=torch.float64, device="cuda")
    Q.fill_momenta(mom, prm, seed, 0, 0)
    prm.close()
    return U, mom


@pytest.mark.parametrize("shapes", [((512, 384), (40, 24)), ((96, 130), (8, 6))])
def test_noalloc_matches_context_entry_points(Q, shapes):
    import torch
    for (L1, L2), beta in zip(shapes, (4.0, 11.25)):
        say(f"noalloc energies {L1}x{L2} beta {beta}")
        U, mom = _fields(Q, torch, L1, L2, 1000 + L1)
        got = _call(Q, torch, L1, L2, beta, mom, U)
        ref = _reference(Q, L1, L2, beta, mom, U)
        assert np.array_equal(got, ref), (got, ref)


def test_noalloc_without_momenta_is_the_action(Q):
    import torch
    L1, L2 = 64, 32
    say(f"noalloc energies {L1}x{L2} without momenta")
    U, mom = _fields(Q, torch, L1, L2, 7)
    got = _call(Q, torch, L1, L2, 2.0, None, U)
    zero = torch.zeros_like(mom)
    ref = _reference(Q, L1, L2, 2.0, zero, U)
    assert np.array_equal(got, ref)
    assert got[0] == got[1]


def test_noalloc_rejects_bad_sizes(Q):
    import torch
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    U = torch.zeros((2, 4, 4), dtype=torch.complex128, device="cuda")
    for L1, L2 in ((1, 4), (4, 1), (1 << 16, 1 << 16)):
        rc = Q.lib.qed2d_hamiltonian_noalloc(L1, L2, C.c_double(1.0), None,
                                             C.c_void_p(U.data_ptr()),
                                             C.c_void_p(out.data_ptr()), None)
        assert rc == Q.QED2D_EINVAL
