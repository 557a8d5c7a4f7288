"""C-ABI checks that need no GPU: libqed2d_hip.so loads, exports every
function include/qed2d.h declares, the Python mirror binds each of them, and
the host-only entry points (Zolotarev coefficients, the Metropolis uniform)
agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qed2d.h")
LIB = os.path.join(ROOT, "dav-qed2d_amd", "qed2d", "libqed2d_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qed2d_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) > 40
    for must in ("qed2d_gamm5Dw", "qed2d_cg", "qed2d_hmc_nf2", "qed2d_hmc_rhmc",
                 "qed2d_tr_dQwdU", "qed2d_krnl_force", "qed2d_updt_fld", "qed2d_zolotarev"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_mirror_binds_every_symbol():
    import qed2d
    assert sorted(qed2d.SIGNATURES) == declared()


def test_zolotarev_host_entry_matches_oracle(O):
    import qed2d
    for n, ra, rb in [(1, 0.3, 5.0), (5, np.sqrt(0.1), np.sqrt(26.0)), (16, 0.07, 4.1)]:
        r = qed2d.get_rhmc_params(n, ra, rb)
        z = O.zolotarev(n, ra, rb)
        assert r.eps == z["eps"]
        np.testing.assert_allclose(r.mu, z["mu"], rtol=1e-15)
        np.testing.assert_allclose(r.nu, z["nu"], rtol=1e-15)
        np.testing.assert_allclose(r.rho, z["rho"], rtol=1e-14)
        assert abs(r.A - z["A"]) <= 1e-15 * abs(z["A"])
        assert abs(r.delta - z["delta"]) <= 1e-14 * abs(z["delta"])


def test_zolotarev_rejects_bad_input():
    import qed2d
    with pytest.raises(qed2d.QED2dError):
        qed2d.get_rhmc_params(0, 0.1, 1.0)
    with pytest.raises(qed2d.QED2dError):
        qed2d.get_rhmc_params(3, 2.0, 1.0)


def test_metropolis_uniform_matches_oracle(O):
    import qed2d
    for t in range(50):
        assert qed2d.metropolis_uniform(1234, 3, t) == O.metropolis_uniform(1234, 3, t)
    assert qed2d.metropolis_uniform(2 ** 40 + 7, 1, 9) == O.metropolis_uniform(2 ** 40 + 7, 1, 9)
