"""qed2d -- Python mirror of the reference Julia module QED2d (dalbandea/Dav-QED2d)
running on libqed2d_hip, the MI355X (gfx950) implementation of its
molecular-dynamics hot path.

Names, argument order and meaning follow the reference (src/QED2d.jl:27-54);
Julia's trailing ``!`` is dropped (``HMC!`` -> ``HMC``, ``plaquette!`` ->
``plaquette``).  Fields are torch tensors on the GPU in the reference's own
layout, Julia column-major ``[L1, L2, 2]``, i.e. C-order shape ``(2, L2, L1)``:

* spinors and links: ``torch.complex128``, shape ``(2, L2, L1)``
* momenta and forces: ``torch.float64``, shape ``(2, L2, L1)``
* per-site densities: ``torch.float64``, shape ``(L2, L1)``

torch only provides device memory and streams; every computation runs in
libqed2d_hip through its C ABI (include/qed2d.h).  There is no CPU fallback:
importing this module without the built library raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import math
import os
import weakref
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libqed2d_hip.so")

if not os.path.exists(LIBPATH):
    raise ImportError(
        f"libqed2d_hip.so not found at {LIBPATH}; build it with "
        "`make -C dav-qed2d_amd/csrc` (or __graft_entry__.build())")

# torch first: the library then binds the HIP runtime torch has loaded, and
# both see the same devices and streams (loaded the other way round, two HIP
# runtimes end up in the process and the library's finds no device).
import torch  # noqa: E402,F401

lib = C.CDLL(LIBPATH)

QED2D_OK, QED2D_EINVAL, QED2D_EHIP, QED2D_ENOCONV, QED2D_ENOMEM = 0, -1, -2, -3, -4
BC_PERIODIC, BC_APERIODIC_T = 0, 1
OP_SQR_MSQ, OP_SQR_MUSQ = 0, 1
PREC_F64, PREC_MIXED = 0, 1

_d, _f, _i, _u32, _u64, _p = C.c_double, C.c_float, C.c_int, C.c_uint32, C.c_uint64, C.c_void_p
_pd, _pi = C.POINTER(C.c_double), C.POINTER(C.c_int)

# Every symbol declared in include/qed2d.h, with its ctypes signature.
SIGNATURES = {
    "qed2d_ctx_create": (_i, [C.POINTER(_p), _i, _i, _d, _i, _p]),
    "qed2d_ctx_destroy": (None, [_p]),
    "qed2d_last_error": (C.c_char_p, [_p]),
    "qed2d_ctx_set_beta": (_i, [_p, _d]),
    "qed2d_ctx_set_sync": (_i, [_p, _i]),
    "qed2d_ctx_set_tile": (_i, [_p, _i]),
    "qed2d_ctx_set_tile_f32": (_i, [_p, _i]),
    "qed2d_ctx_set_tiles": (_i, [_p, _i, _i, _i, _i]),
    "qed2d_ctx_get_tiles": (_i, [_p, _p]),
    "qed2d_ctx_set_tiles_k2": (_i, [_p, _i, _i]),
    "qed2d_ctx_get_tiles_k2": (_i, [_p, _p]),
    "qed2d_batch_create": (_i, [_p, _i, _p, _p, _p, _i, _p]),
    "qed2d_batch_destroy": (None, [_p]),
    "qed2d_batch_last_error": (C.c_char_p, [_p]),
    "qed2d_batch_size": (_i, [_p, _p, _p]),
    "qed2d_batch_offset": (_i, [_p, _i, _p]),
    "qed2d_batch_links": (_i, [_p, _i, _p]),
    "qed2d_batch_get_links": (_i, [_p, _i, _p]),
    "qed2d_batch_set_links": (_i, [_p, _i, _p]),
    "qed2d_batch_set_tiles": (_i, [_p, _i, _i, _i]),
    "qed2d_batch_set_tiles_k2": (_i, [_p, _i, _i]),
    "qed2d_batch_get_tiles": (_i, [_p, _p]),
    "qed2d_batch_autotune": (_i, [_p, _i, _p]),
    "qed2d_batch_set_graphs": (_i, [_p, _i]),
    "qed2d_batch_set_reliable_mode": (_i, [_p, _i]),
    "qed2d_batch_hot_start": (_i, [_p, _p]),
    "qed2d_batch_cg": (_i, [_p, _p, _p, _p, _i, _d, _i, _p]),
    "qed2d_batch_create_rhs": (_i, [_p, _i, _i, _i, _d, _i, _p]),
    "qed2d_batch_cg_shared": (_i, [_p, _p, _p, _p, _p, _i, _d, _i, _p]),
    "qed2d_batch_hmc_nf2": (_i, [_p, _p, _d, _i, _i, _i, _d, _i, _d, _p, _u32, _p, _p, _p, _p]),
    "qed2d_batch_observables": (_i, [_p, _p, _p, _p]),
    "qed2d_batch_last_dh": (_i, [_p, _p]),
    "qed2d_ctx_last_dh": (_i, [_p, _p]),
    "qed2d_hamiltonian_noalloc": (_i, [_i, _i, _d, _p, _p, _p, _p]),
    "qed2d_ctx_autotune": (_i, [_p, _i, _p]),
    "qed2d_hbm_stream": (_i, [_p, C.c_long, _p]),
    "qed2d_ctx_set_reliable_delta": (_i, [_p, _d]),
    "qed2d_ctx_set_reliable_mode": (_i, [_p, _i]),
    "qed2d_ctx_set_rhmc_precision": (_i, [_p, _i]),
    "qed2d_synchronize": (_i, [_p]),
    "qed2d_gamm5Dw": (_i, [_p, _p, _p, _p, _d]),
    "qed2d_gamm5Dw_f32": (_i, [_p, _p, _p, _p, _f]),
    "qed2d_gamm5Dw_mrhs": (_i, [_p, _p, _p, _p, _i, _d]),
    "qed2d_gamm5Dw_mrhs_f32": (_i, [_p, _p, _p, _p, _i, _f]),
    "qed2d_gamm5": (_i, [_p, _p, _p]),
    "qed2d_Dw": (_i, [_p, _p, _p, _p, _d]),
    "qed2d_gamm5Dw_sqr": (_i, [_p, _p, _p, _p, _d]),
    "qed2d_gamm5Dw_sqr_msq": (_i, [_p, _p, _p, _p, _p, _d]),
    "qed2d_gamm5Dw_sqr_musq": (_i, [_p, _p, _p, _p, _p, _d, _d]),
    "qed2d_gamm5Dw_sqr_sqr": (_i, [_p, _p, _p, _p, _d]),
    "qed2d_gamm5Dw_sqr_sqr_musq": (_i, [_p, _p, _p, _p, _p, _d, _d]),
    "qed2d_tr_dQwdU": (_i, [_p, _p, _p, _p, _p]),
    "qed2d_krnl_force": (_i, [_p, _p, _p, _p]),
    "qed2d_updt_fld": (_i, [_p, _p, _p, _d]),
    "qed2d_force_mom_update": (_i, [_p, _p, _p, _p, _p, _d]),
    "qed2d_plaquette_field": (_i, [_p, _p, _p]),
    "qed2d_qtop_field": (_i, [_p, _p, _p]),
    "qed2d_action": (_i, [_p, _p, _pd]),
    "qed2d_plaquette": (_i, [_p, _p, _pd]),
    "qed2d_qtop": (_i, [_p, _p, _pd]),
    "qed2d_hamiltonian": (_i, [_p, _p, _p, _pd]),
    "qed2d_dot": (_i, [_p, _p, _p, _pd, _pd]),
    "qed2d_norm2": (_i, [_p, _p, _pd]),
    "qed2d_cg": (_i, [_p, _p, _p, _p, _d, _d, _i, _i, _d, _i, _i, _pi]),
    "qed2d_cg_mrhs": (_i, [_p, _p, _p, _p, _i, _d, _i, _d, _i, _pi]),
    "qed2d_multicg": (_i, [_p, _p, _p, _p, _d, _i, _pd, _pd, _i, _d]),
    "qed2d_R": (_i, [_p, _p, _p, _p, _d, _i, _pd, _pd, _d, _d, _i, _d]),
    "qed2d_generate_pseudofermion": (_i, [_p, _p, _p, _p, _d, _i, _pd, _pd, _i, _d]),
    "qed2d_luscher_z": (_i, [_p, _p, _i, _p, _p, _d, _i, _pd, _pd, _d, _d, _i, _d]),
    "qed2d_reweighting_factor": (_i, [_p, _p, _d, _i, _pd, _pd, _d, _d, _i, _i, _i, _d, _u64,
                                      _u32, _u32, _pd]),
    "qed2d_power_method": (_i, [_p, _p, _p, _d, _i, _pd, _pd]),
    "qed2d_zolotarev": (_i, [_i, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd]),
    "qed2d_fill_momenta": (_i, [_p, _p, _u64, _u32, _u32]),
    "qed2d_fill_pseudofermion": (_i, [_p, _p, _u64, _u32, _u32, _u32]),
    "qed2d_hot_start": (_i, [_p, _p, _u64, _u32]),
    "qed2d_metropolis_uniform": (_d, [_u64, _u32, _u32]),
    "qed2d_omf4_pure": (_i, [_p, _p, _p, _d, _i]),
    "qed2d_md_nf2": (_i, [_p, _i, _p, _p, _p, _p, _p, _d, _d, _i, _i, _d, _i, _pi]),
    "qed2d_hmc_pure": (_i, [_p, _p, _d, _i, _i, _i, _u64, _u32, _u32, _pd, _pd, _pi]),
    "qed2d_hmc_nf2": (_i, [_p, _p, _d, _d, _i, _i, _i, _d, _i, _i, _d, _u64, _u32, _u32, _pd,
                           _pd, _pi, _pi]),
    "qed2d_hmc_rhmc": (_i, [_p, _p, _pd, _i, _pi, _p, _p, _p, _d, _i, _i, _i, _d, _i, _i, _u64,
                            _u32, _u32, _pd, _pd, _pi]),
    "qed2d_unfold": (_i, [_p, _p, _p, _i]),
    "qed2d_multishift_cg": (_i, [_p, C.POINTER(_p), _p, _p, _d, _i, _pd, _i, _d, _pi]),
    "qed2d_ctx_set_multishift": (_i, [_p, _i]),
    "qed2d_save_gauge": (_i, [_p, _p, C.c_char_p]),
    "qed2d_load_gauge": (_i, [_p, _p, C.c_char_p]),
    "qed2d_write_links_text": (_i, [_p, C.c_long, C.c_char_p]),
    "qed2d_read_links_text": (_i, [_p, C.c_long, C.c_char_p]),
    "qed2d_profile_enable": (_i, [_p, _i]),
    "qed2d_profile_reset": (_i, [_p]),
    "qed2d_profile_read": (_i, [_p, _i, C.POINTER(C.c_long), _pd]),
}
for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


PROF_CLASSES = ["cg_k1_f64", "cg_k2_f64", "cg_k1_f32", "cg_k2_f32", "q_f64", "force_mom",
                "updt_fld", "reliable_update", "other"]


def set_tiles(prm, q64=0, q32=0, cg64=0, cg32=0, cg64_k2=None, cg32_k2=None):
    """Tile heights (rows marched per thread) of the FP64 / FP32 Q applies and
    the FP64 / FP32 solver kernels; 0 leaves a family unchanged.  A nonzero
    cg64 / cg32 also sets the second CG kernel's height, unless cg64_k2 /
    cg32_k2 give it separately."""
    _check(prm, lib.qed2d_ctx_set_tiles(prm.ctx(), int(q64), int(q32), int(cg64), int(cg32)))
    if cg64_k2 is not None or cg32_k2 is not None:
        cur = get_tiles(prm)
        a = cur["cg64"] if cg64_k2 is None else cg64_k2
        b = cur["cg32"] if cg32_k2 is None else cg32_k2
        _check(prm, lib.qed2d_ctx_set_tiles_k2(prm.ctx(), int(a), int(b)))


def get_tiles(prm):
    """The tile heights in use: the four families, and the second CG kernel's
    (cg64_k2, cg32_k2)."""
    t = (_i * 4)()
    _check(prm, lib.qed2d_ctx_get_tiles(prm.ctx(), t))
    t2 = (_i * 2)()
    _check(prm, lib.qed2d_ctx_get_tiles_k2(prm.ctx(), t2))
    return dict(zip(("q64", "q32", "cg64", "cg32", "cg64_k2", "cg32_k2"), list(t) + list(t2)))


def autotune(prm, iters=160):
    """Times every tile height of each kernel family on this lattice and keeps
    the fastest (returns the choice, as get_tiles); the CG families over
    `iters` iterations, every pair of the two kernels' heights in two passes."""
    t = (_i * 4)()
    _check(prm, lib.qed2d_ctx_autotune(prm.ctx(), int(iters), t))
    return get_tiles(prm)


def hbm_stream(prm, nbytes=1 << 30):
    """STREAM copy / triad / read bandwidth of the device in GB/s, and the copy
    of the same bytes as 16 arrays into 16 others (copy16)."""
    g = (_d * 4)()
    _check(prm, lib.qed2d_hbm_stream(prm.ctx(), int(nbytes), g))
    return dict(zip(("copy", "triad", "read", "copy16"), list(g)))


def profile_enable(prm, on=True, every=1):
    """Per-kernel-class HIP-event timing on the context's stream; `every`
    samples one launch in `every` of each class."""
    _check(prm, lib.qed2d_profile_enable(prm.ctx(), int(every) if on else 0))
    _check(prm, lib.qed2d_profile_reset(prm.ctx()))


def profile_read(prm):
    """{class: (launches, total_ms)} since profile_enable."""
    out = {}
    for k, name in enumerate(PROF_CLASSES):
        n, ms = C.c_long(), _d()
        _check(prm, lib.qed2d_profile_read(prm.ctx(), k, C.byref(n), C.byref(ms)))
        out[name] = (n.value, ms.value)
    return out


class QED2dError(RuntimeError):
    """Raised for a nonzero status of libqed2d_hip (the reference throws)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class CGNotConverged(QED2dError):
    pass


# ---------------------------------------------------------------------------
# parameter types (src/QED2d.jl:5-26, src/integrators/integrators.jl:1-3)
# ---------------------------------------------------------------------------
_live = weakref.WeakSet()
_exiting = False


@atexit.register
def _destroy_all():
    """Release every context before the HIP runtime tears itself down."""
    global _exiting
    for prm in list(_live):
        prm.close()
    _exiting = True


class LattParm:
    """LattParm(iL, beta) (src/QED2d.jl:5-8); ``bc`` adds the fermion boundary
    flag (0 periodic as the reference, BC_APERIODIC_T antiperiodic in i2)."""

    def __init__(self, iL, beta, bc=BC_PERIODIC):
        self.iL = (int(iL[0]), int(iL[1]))
        self.beta = float(beta)
        self.bc = int(bc)
        self._ctx = None
        self._ctx_beta = None

    @property
    def V(self):
        return self.iL[0] * self.iL[1]

    def ctx(self):
        import torch
        if self._ctx is None:
            h = _p()
            stream = torch.cuda.current_stream().cuda_stream
            rc = lib.qed2d_ctx_create(C.byref(h), self.iL[0], self.iL[1], self.beta, self.bc,
                                      _p(stream))
            if rc != QED2D_OK:
                raise QED2dError(rc, "qed2d_ctx_create failed: " +
                                 lib.qed2d_last_error(None).decode())
            self._ctx = h
            self._ctx_beta = self.beta
            _live.add(self)
        if self._ctx_beta != self.beta:
            lib.qed2d_ctx_set_beta(self._ctx, self.beta)
            self._ctx_beta = self.beta
        return self._ctx

    def close(self):
        if getattr(self, "_ctx", None) is not None and not _exiting:
            lib.qed2d_ctx_destroy(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class KernelParm:
    """KernelParm(threads, blocks) (src/QED2d.jl:10-13).  Accepted for
    signature compatibility; launch shapes are chosen by the library."""
    threads: tuple = (1, 1)
    blocks: tuple = (1, 1)


@dataclass
class RHMCParm:
    """RHMCParm (src/QED2d.jl:15-26)."""
    r_b: float
    n: int
    eps: float
    A: float
    rho: List[float]
    mu: List[float]
    nu: List[float]
    delta: float
    reweighting_N: int = 1
    reweighting_Taylor: int = 5


class Integrators:
    pass


class Leapfrog(Integrators):
    code = 0


class OMF4(Integrators):
    code = 1


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _check(prm, rc):
    if rc != QED2D_OK:
        msg = lib.qed2d_last_error(prm._ctx).decode() if prm._ctx is not None else ""
        if rc == QED2D_ENOCONV:
            raise CGNotConverged(rc, msg)
        raise QED2dError(rc, msg)


def _ptr(t, dtype=None, shape=None):
    import torch
    if t is None:
        return None
    if not t.is_cuda:
        raise TypeError("fields must be CUDA (HIP) tensors")
    if not t.is_contiguous():
        raise ValueError("fields must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return _p(t.data_ptr())


def _sp(prm, t):
    import torch
    return _ptr(t, torch.complex128, (2, prm.iL[1], prm.iL[0]))


def _sp32(prm, t):
    import torch
    return _ptr(t, torch.complex64, (2, prm.iL[1], prm.iL[0]))


def _re(prm, t):
    import torch
    return _ptr(t, torch.float64, (2, prm.iL[1], prm.iL[0]))


def _site(prm, t):
    import torch
    return _ptr(t, torch.float64, (prm.iL[1], prm.iL[0]))


def spinor(prm, device="cuda"):
    import torch
    return torch.zeros((2, prm.iL[1], prm.iL[0]), dtype=torch.complex128, device=device)


def realfield(prm, device="cuda"):
    import torch
    return torch.zeros((2, prm.iL[1], prm.iL[0]), dtype=torch.float64, device=device)


def cold_links(prm, device="cuda"):
    """U = 1 everywhere (CUDA.ones(ComplexF64, L1, L2, 2))."""
    import torch
    return torch.ones((2, prm.iL[1], prm.iL[0]), dtype=torch.complex128, device=device)


# ---------------------------------------------------------------------------
# src/QED2dDirac.jl
# ---------------------------------------------------------------------------
def gamm5Dw(so, U, si, am0, prm):
    """Kernel gamm5Dw (src/QED2dDirac.jl:2-33): so = gamma5 Dw si."""
    _check(prm, lib.qed2d_gamm5Dw(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si), am0))


def gamm5Dw_f32(so, U, si, am0, prm):
    """FP32 form of gamm5Dw (complex64 fields)."""
    _check(prm, lib.qed2d_gamm5Dw_f32(prm.ctx(), _sp32(prm, so), _sp32(prm, U), _sp32(prm, si),
                                      am0))


def gamm5Dw_mrhs(so, U, si, am0, prm):
    """gamm5Dw on k right-hand sides sharing U; si, so shaped (k, 2, L2, L1)."""
    import torch
    k = si.shape[0]
    if so.shape != si.shape or si.shape[1:] != (2, prm.iL[1], prm.iL[0]):
        raise ValueError("multi-RHS fields must be (k, 2, L2, L1)")
    if si.dtype == torch.complex128:
        _check(prm, lib.qed2d_gamm5Dw_mrhs(prm.ctx(), _ptr(so), _sp(prm, U), _ptr(si), k, am0))
    else:
        _check(prm, lib.qed2d_gamm5Dw_mrhs_f32(prm.ctx(), _ptr(so), _sp32(prm, U), _ptr(si), k,
                                               am0))


def gamm5(so, si, am0, prm):
    """Kernel gamm5 (src/QED2dDirac.jl:130-146)."""
    _check(prm, lib.qed2d_gamm5(prm.ctx(), _sp(prm, so), _sp(prm, si)))


def Dw(so, U, si, am0, prm, kprm=None):
    """Dw (src/QED2dDirac.jl:157-169)."""
    _check(prm, lib.qed2d_Dw(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si), am0))


def gamm5Dw_sqr(so, U, si, am0, prm, kprm=None):
    """(gamma5 Dw)^2 (src/QED2dDirac.jl:35-46)."""
    _check(prm, lib.qed2d_gamm5Dw_sqr(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si), am0))


def gamm5Dw_sqr_sqr(so, U, si, am0, prm, kprm=None):
    """(gamma5 Dw)^4 (src/QED2dDirac.jl:48-67)."""
    _check(prm, lib.qed2d_gamm5Dw_sqr_sqr(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si),
                                          am0))


def gamm5Dw_sqr_msq(so, tmp, U, si, am0, prm, kprm=None):
    """(gamma5 Dw)^2, the Nf=2 CG operator (src/QED2dDirac.jl:69-81)."""
    _check(prm, lib.qed2d_gamm5Dw_sqr_msq(prm.ctx(), _sp(prm, so), _sp(prm, tmp), _sp(prm, U),
                                          _sp(prm, si), am0))


def gamm5Dw_sqr_musq(so, tmp, U, si, am0, mu_j, prm, kprm=None):
    """(gamma5 Dw)^2 + mu^2, the RHMC operator (src/QED2dDirac.jl:83-95)."""
    _check(prm, lib.qed2d_gamm5Dw_sqr_musq(prm.ctx(), _sp(prm, so), _sp(prm, tmp), _sp(prm, U),
                                           _sp(prm, si), am0, mu_j))


def gamm5Dw_sqr_sqr_musq(so, tmp, U, si, am0, mu_j, prm, kprm=None):
    """(gamma5 Dw)^4 + mu^2 (src/QED2dDirac.jl:97-117)."""
    _check(prm, lib.qed2d_gamm5Dw_sqr_sqr_musq(prm.ctx(), _sp(prm, so), _sp(prm, tmp),
                                               _sp(prm, U), _sp(prm, si), am0, mu_j))


def tr_dQwdU(frc, U, X, g5DwX, prm):
    """Kernel tr_dQwdU, the fermion force (src/QED2dDirac.jl:172-199)."""
    _check(prm, lib.qed2d_tr_dQwdU(prm.ctx(), _re(prm, frc), _sp(prm, U), _sp(prm, X),
                                   _sp(prm, g5DwX)))


# ---------------------------------------------------------------------------
# src/QED2dAction.jl
# ---------------------------------------------------------------------------
def plaquette(plx, U, prm):
    """Kernel plaquette! (src/QED2dAction.jl:2-16)."""
    _check(prm, lib.qed2d_plaquette_field(prm.ctx(), _site(prm, plx), _sp(prm, U)))


def qtop(plx, U, prm):
    """Kernel qtop! (src/QED2dAction.jl:18-32)."""
    _check(prm, lib.qed2d_qtop_field(prm.ctx(), _site(prm, plx), _sp(prm, U)))


def krnl_force(frc1, frc2, U, prm):
    """Kernel krnl_force! (src/QED2dAction.jl:34-53)."""
    _check(prm, lib.qed2d_krnl_force(prm.ctx(), _re(prm, frc1), _re(prm, frc2), _sp(prm, U)))


def updt_fld(U, mom, eps, prm):
    """Kernel updt_fld! (src/QED2dAction.jl:55-64).  The reference kernel takes
    no LattParm; the mirror needs it to find the lattice context."""
    _check(prm, lib.qed2d_updt_fld(prm.ctx(), _sp(prm, U), _re(prm, mom), eps))


def force_mom_update(mom, U, X, g5DwX, eps, prm):
    """mom .+= eps (frc1 + frc2 + frc) with the forces fused (src/QED2dHMC.jl:463-478)."""
    _check(prm, lib.qed2d_force_mom_update(prm.ctx(), _re(prm, mom), _sp(prm, U),
                                           _sp(prm, X) if X is not None else None,
                                           _sp(prm, g5DwX) if g5DwX is not None else None, eps))


# ---------------------------------------------------------------------------
# src/QED2dHMC.jl observables
# ---------------------------------------------------------------------------
def _scalar(prm, fn, *args):
    out = _d()
    _check(prm, fn(prm.ctx(), *args, C.byref(out)))
    return out.value


def Action(U, prm, kprm=None):
    """Action (src/QED2dHMC.jl:206-215)."""
    return _scalar(prm, lib.qed2d_action, _sp(prm, U))


def Plaquette(U, prm, kprm=None):
    """Plaquette (src/QED2dHMC.jl:217-226)."""
    return _scalar(prm, lib.qed2d_plaquette, _sp(prm, U))


def Qtop(U, prm, kprm=None):
    """Qtop (src/QED2dHMC.jl:228-237)."""
    return _scalar(prm, lib.qed2d_qtop, _sp(prm, U))


def Hamiltonian(mom, U, prm, kprm=None):
    """Hamiltonian (src/QED2dHMC.jl:198-203)."""
    return _scalar(prm, lib.qed2d_hamiltonian, _re(prm, mom), _sp(prm, U))


def dot(a, b, prm):
    """CUDA.dot(a, b) = sum conj(a) b on spinors."""
    re, im = _d(), _d()
    _check(prm, lib.qed2d_dot(prm.ctx(), _sp(prm, a), _sp(prm, b), C.byref(re), C.byref(im)))
    return complex(re.value, im.value)


def norm2(a, prm):
    """CUDA.mapreduce(abs2, +, a)."""
    return _scalar(prm, lib.qed2d_norm2, _sp(prm, a))


# ---------------------------------------------------------------------------
# src/QED2dSolver.jl
# ---------------------------------------------------------------------------
_OPS = {gamm5Dw_sqr_msq: OP_SQR_MSQ, gamm5Dw_sqr_musq: OP_SQR_MUSQ}


def CG(*args, so_as_guess=False, precision=PREC_F64):
    """CG (src/QED2dSolver.jl).  Two methods, as the reference:

    * ``CG(so, U, si, am0, maxiter, eps, A, prm, kprm)``      (:3-45)
    * ``CG(so, U, si, am0, mu, maxiter, eps, A, prm, kprm; so_as_guess)``  (:49-105)

    ``A`` must be ``gamm5Dw_sqr_msq`` or ``gamm5Dw_sqr_musq``.  ``eps`` is a
    squared relative residual.  Returns the iteration count; raises
    ``CGNotConverged`` after ``maxiter`` iterations.  ``precision=PREC_MIXED``
    runs FP32 iterations with FP64 reliable updates to the same FP64 stopping
    rule.
    """
    if len(args) == 9:
        so, U, si, am0, maxiter, eps, A, prm, kprm = args
        mu = 0.0
    elif len(args) == 10:
        so, U, si, am0, mu, maxiter, eps, A, prm, kprm = args
    else:
        raise TypeError("CG takes 9 or 10 positional arguments")
    if A not in _OPS:
        raise TypeError("A must be gamm5Dw_sqr_msq or gamm5Dw_sqr_musq")
    it = _i()
    _check(prm, lib.qed2d_cg(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si), am0, mu,
                             _OPS[A], int(maxiter), eps, 1 if so_as_guess else 0, precision,
                             C.byref(it)))
    return it.value


def CG_mrhs(so, U, si, am0, maxiter, eps, prm, kprm=None, precision=PREC_F64):
    """CG (src/QED2dSolver.jl:3-45, operator ``gamm5Dw_sqr_msq``) on k
    right-hand sides sharing the links, all of them in one launch per kernel.
    ``si`` and ``so`` have shape ``(k, 2, L2, L1)``; ``so[j]`` is bit for bit
    what ``CG(so[j], U, si[j], ...)`` returns with the same tiles.  Returns the
    iteration counts; raises ``CGNotConverged`` if any right-hand side does not
    converge."""
    import torch
    k = si.shape[0]
    shp = (k, 2, prm.iL[1], prm.iL[0])
    if tuple(si.shape) != shp or tuple(so.shape) != shp:
        raise ValueError(f"so and si must have shape {shp}")
    its = (_i * k)()
    _check(prm, lib.qed2d_cg_mrhs(prm.ctx(), _ptr(so, torch.complex128, shp), _sp(prm, U),
                                  _ptr(si, torch.complex128, shp), k, am0, int(maxiter), eps,
                                  precision, its))
    return list(its)


# ---------------------------------------------------------------------------
# src/QED2dRHMC.jl
# ---------------------------------------------------------------------------
def get_rhmc_params(n_rhmc, r_a_rhmc, r_b_rhmc, reweighting_N=1, reweighting_Taylor=5):
    """get_rhmc_params (src/QED2dRHMC.jl:93-121): scalar or list arguments."""
    if isinstance(n_rhmc, (list, tuple, np.ndarray)):
        return [get_rhmc_params(int(n), a, b) for n, a, b in zip(n_rhmc, r_a_rhmc, r_b_rhmc)]
    n = int(n_rhmc)
    eps, A, delta = _d(), _d(), _d()
    rho, mu, nu = (_d * n)(), (_d * n)(), (_d * n)()
    rc = lib.qed2d_zolotarev(n, float(r_a_rhmc), float(r_b_rhmc), C.byref(eps), C.byref(A),
                             C.byref(delta), rho, mu, nu)
    if rc != QED2D_OK:
        raise QED2dError(rc, f"invalid Zolotarev parameters n={n} r_a={r_a_rhmc} r_b={r_b_rhmc}")
    return RHMCParm(float(r_b_rhmc), n, eps.value, A.value, list(rho), list(mu), list(nu),
                    delta.value, reweighting_N, reweighting_Taylor)


def _arr(v):
    a = (_d * len(v))(*[float(x) for x in v])
    return a


def multishift_CG(xs, U, si, am0, mus, maxiter, eps, prm):
    """(Q^2 + mu_i^2) x_i = si for every shift at once (multi-shift CG).
    Returns the iteration count of the slowest shift."""
    n = len(mus)
    ptrs = (_p * n)(*[_sp(prm, x).value for x in xs])
    it = _i()
    _check(prm, lib.qed2d_multishift_cg(prm.ctx(), ptrs, _sp(prm, U), _sp(prm, si), am0, n,
                                        _arr(mus), int(maxiter), eps, C.byref(it)))
    return it.value


def set_reliable_mode(prm, on_device):
    """Mixed-precision CG: reliable updates decided on the device (True, the
    default; the host never waits for one) or by the host after a pause
    (False: the batched solvers' schedule, which CG_mrhs and Batch reproduce
    bit for bit)."""
    _check(prm, lib.qed2d_ctx_set_reliable_mode(prm.ctx(), 1 if on_device else 0))


def set_rhmc_precision(prm, precision):
    """Precision of the RHMC force's solves (PREC_F64, the default, or
    PREC_MIXED: FP32 multi-shift, then each pole from that guess by a batched
    mixed-precision CG to the FP64 stopping rule); see
    qed2d_ctx_set_rhmc_precision.  No counterpart in the reference."""
    _check(prm, lib.qed2d_ctx_set_rhmc_precision(prm.ctx(), int(precision)))
    prm._rhmc_prec = int(precision)


def set_multishift(prm, on):
    """MultiCG / R / RHMC force through multi-shift CG (True or 1: all
    pseudofermions of the RHMC force in one solve; 2: one pseudofermion at a
    time) or through the reference's sequential CG solves (False or 0)."""
    mode = int(on) if not isinstance(on, bool) else (1 if on else 0)
    _check(prm, lib.qed2d_ctx_set_multishift(prm.ctx(), mode))


def MultiCG(so, U, si, am0, maxiter, eps, A, rprm, prm, kprm=None):
    """MultiCG (src/QED2dRHMC.jl:281-290)."""
    _check(prm, lib.qed2d_multicg(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si), am0,
                                  rprm.n, _arr(rprm.mu), _arr(rprm.rho), int(maxiter), eps))


def R(so, U, si, am0, maxiter, eps, A, rprm, prm, kprm=None):
    """R (src/QED2dRHMC.jl:261-271)."""
    _check(prm, lib.qed2d_R(prm.ctx(), _sp(prm, so), _sp(prm, U), _sp(prm, si), am0, rprm.n,
                            _arr(rprm.mu), _arr(rprm.rho), rprm.A, rprm.r_b, int(maxiter), eps))


def generate_pseudofermion(F, U, X, am0, CGmaxiter, CGtol, prm, kprm, rprm):
    """generate_pseudofermion! (src/QED2dRHMC.jl:298-319)."""
    _check(prm, lib.qed2d_generate_pseudofermion(prm.ctx(), _sp(prm, F), _sp(prm, U),
                                                 _sp(prm, X), am0, rprm.n, _arr(rprm.mu),
                                                 _arr(rprm.nu), int(CGmaxiter), CGtol))


def LuscherZp(Z, p, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm=None):
    """LuscherZp (src/QED2dRHMC.jl:189-201): Z = (Q^2 R^2 - 1)^p X_in."""
    _check(prm, lib.qed2d_luscher_z(prm.ctx(), _sp(prm, Z), int(p), _sp(prm, U), _sp(prm, X_in),
                                    am0, rprm.n, _arr(rprm.mu), _arr(rprm.rho), rprm.A, rprm.r_b,
                                    int(CGmaxiter), CGtol))


def LuscherZ(Z, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm=None):
    """LuscherZ (src/QED2dRHMC.jl:174-186): Z = (Q^2 R^2 - 1) X_in."""
    LuscherZp(Z, 1, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm)


def reweighting_factor(U, am0, CGmaxiter, CGtol, prm, kprm, rprm):
    """reweighting_factor (src/QED2dRHMC.jl:210-251).  With lists of masses and
    RHMCParm, the product over fermions.  The Gaussian fields come from the
    build's stream (seed(); one trajectory index per fermion)."""
    if isinstance(rprm, (list, tuple)):
        w = 1.0
        for m, r in zip(am0, rprm):
            w *= reweighting_factor(U, m, CGmaxiter, CGtol, prm, kprm, r)
        return w
    W = _d()
    _check(prm, lib.qed2d_reweighting_factor(prm.ctx(), _sp(prm, U), am0, rprm.n, _arr(rprm.mu),
                                             _arr(rprm.rho), rprm.A, rprm.r_b,
                                             int(rprm.reweighting_N),
                                             int(rprm.reweighting_Taylor), int(CGmaxiter), CGtol,
                                             _rng.seed, _rng.chain, _rng.next_traj(),
                                             C.byref(W)))
    return W.value


def power_method(U, am0, prm, kprm=None, iter=1000, b0=None):
    """power_method (src/QED2dRHMC.jl:135-170).  The starting vector is drawn
    from this package's own stream (seed(), then flavour 15) unless given."""
    if b0 is None:
        b0 = spinor(prm)
        _check(prm, lib.qed2d_fill_pseudofermion(prm.ctx(), _sp(prm, b0), _rng.seed, _rng.chain,
                                                 _rng.next_traj(), 15))
    lmin, lmax = _d(), _d()
    _check(prm, lib.qed2d_power_method(prm.ctx(), _sp(prm, U), _sp(prm, b0), am0, int(iter),
                                       C.byref(lmin), C.byref(lmax)))
    return lmin.value, lmax.value


# ---------------------------------------------------------------------------
# random streams: Random.seed!(CURAND.default_rng(), seed); Random.seed!(seed)
# ---------------------------------------------------------------------------
class _Stream:
    def __init__(self):
        self.seed = 1234
        self.chain = 0
        self.traj = 0

    def next_traj(self):
        t = self.traj
        self.traj += 1
        return t


_rng = _Stream()


def seed(s, chain=0, traj=0):
    """Counterpart of Random.seed!: sets this package's counter-based stream."""
    _rng.seed = int(s)
    _rng.chain = int(chain)
    _rng.traj = int(traj)


def rng_state():
    return _rng.seed, _rng.chain, _rng.traj


def fill_momenta(mom, prm, seed, chain, traj):
    _check(prm, lib.qed2d_fill_momenta(prm.ctx(), _re(prm, mom), seed, chain, traj))


def fill_pseudofermion(X, prm, seed, chain, traj, flavour=0):
    _check(prm, lib.qed2d_fill_pseudofermion(prm.ctx(), _sp(prm, X), seed, chain, traj,
                                             flavour))


def hot_start(U, prm, seed, chain=0):
    _check(prm, lib.qed2d_hot_start(prm.ctx(), _sp(prm, U), seed, chain))


def metropolis_uniform(seed, chain, traj):
    return lib.qed2d_metropolis_uniform(seed, chain, traj)


# ---------------------------------------------------------------------------
# MD and HMC (src/QED2dHMC.jl)
# ---------------------------------------------------------------------------
def OMF4_pure(mom, U, eps, ns, prm, kprm=None):
    """OMF4! for the pure gauge theory (src/QED2dHMC.jl:239-305)."""
    _check(prm, lib.qed2d_omf4_pure(prm.ctx(), _re(prm, mom), _sp(prm, U), eps, int(ns)))


def MD_nf2(integrator, mom, U, X, F, g5DX, am0, eps, ns, CGmaxiter, tol, prm, kprm=None,
           precision=PREC_F64):
    """MD! for Nf=2 (src/QED2dHMC.jl:5-11): leapfrog! (:426-449) or OMF4! (:307-422).
    Returns the total number of CG iterations."""
    it = _i()
    _check(prm, lib.qed2d_md_nf2(prm.ctx(), integrator.code, _re(prm, mom), _sp(prm, U),
                                 _sp(prm, X), _sp(prm, F), _sp(prm, g5DX), am0, eps, int(ns),
                                 int(CGmaxiter), tol, precision, C.byref(it)))
    return it.value


@dataclass
class TrajInfo:
    hini: float
    hfin: float
    accepted: int
    cg_iters: int = 0
    traj: int = 0
    dh: float = None  # hfin - hini from the double-double energies (below the ulp of H)

    @property
    def dH(self):
        return self.hfin - self.hini if self.dh is None else self.dh


def HMC(U, *args, integrator=None, qzero=False, reversibility_test=False,
        precision=None, acc_tol=0.0, traj=None):
    """HMC! (src/QED2dHMC.jl).  Three methods, as the reference:

    * pure gauge ``HMC(U, eps, ns, acc, prm, kprm)``                  (:19-60)
    * Nf=2       ``HMC(U, am0, eps, ns, acc, CGmaxiter, tol, prm, kprm)``   (:64-113)
    * RHMC       ``HMC(U, am0::list, eps, ns, acc, CGmaxiter, tol, prm, kprm, rprm::list)``
                                                                      (:116-195)

    ``acc`` is appended with 1/0 (accept/reject) as the reference's
    ``push!(acc, ...)``.  Random numbers come from the stream set by
    ``seed()``; each call consumes one trajectory index.  Returns TrajInfo.
    ``precision`` selects the MD force's solves: Nf=2 (default PREC_F64) and,
    for this call, RHMC (default: the context's, ``set_rhmc_precision``).
    """
    t = _rng.next_traj() if traj is None else int(traj)
    hi, hf, acc_o, it = _d(), _d(), _i(), _i()
    if len(args) == 5:
        eps, ns, acc, prm, kprm = args
        integ = integrator or OMF4()
        _check(prm, lib.qed2d_hmc_pure(prm.ctx(), _sp(prm, U), eps, int(ns), integ.code,
                                       1 if qzero else 0, _rng.seed, _rng.chain, t, C.byref(hi),
                                       C.byref(hf), C.byref(acc_o)))
    elif len(args) == 8:
        am0, eps, ns, acc, CGmaxiter, tol, prm, kprm = args
        integ = integrator or Leapfrog()
        prec = PREC_F64 if precision is None else precision
        _check(prm, lib.qed2d_hmc_nf2(prm.ctx(), _sp(prm, U), am0, eps, int(ns), integ.code,
                                      int(CGmaxiter), tol, 1 if qzero else 0, prec, acc_tol,
                                      _rng.seed, _rng.chain, t, C.byref(hi), C.byref(hf),
                                      C.byref(acc_o), C.byref(it)))
    elif len(args) == 9:
        am0s, eps, ns, acc, CGmaxiter, tol, prm, kprm, rprms = args
        # OMF4 is an extension: the reference's MD! has no OMF4 method for RHMC.
        integ = integrator or Leapfrog()
        nf = len(am0s)
        keep = []
        mus = (_p * nf)()
        nus = (_p * nf)()
        rhos = (_p * nf)()
        for j, r in enumerate(rprms):
            a, b, c = _arr(r.mu), _arr(r.nu), _arr(r.rho)
            keep += [a, b, c]
            mus[j] = C.cast(a, _p)
            nus[j] = C.cast(b, _p)
            rhos[j] = C.cast(c, _p)
        npole = (_i * nf)(*[r.n for r in rprms])
        keep_prec = getattr(prm, "_rhmc_prec", PREC_F64)
        if precision is not None:
            _check(prm, lib.qed2d_ctx_set_rhmc_precision(prm.ctx(), int(precision)))
        try:
            _check(prm, lib.qed2d_hmc_rhmc(prm.ctx(), _sp(prm, U), _arr(am0s), nf, npole, mus,
                                           nus, rhos, eps, int(ns), integ.code, int(CGmaxiter),
                                           tol, 1 if qzero else 0, 1 if reversibility_test else 0,
                                           _rng.seed, _rng.chain, t, C.byref(hi), C.byref(hf),
                                           C.byref(acc_o)))
        finally:
            if precision is not None:
                lib.qed2d_ctx_set_rhmc_precision(prm.ctx(), int(keep_prec))
    else:
        raise TypeError("HMC takes 5 (pure gauge), 8 (Nf=2) or 9 (RHMC) positional arguments "
                        "after U")
    if acc_o.value >= 0 and acc is not None:
        acc.append(acc_o.value)
    dh = _d()
    lib.qed2d_ctx_last_dh(prm.ctx(), C.byref(dh))
    return TrajInfo(hi.value, hf.value, acc_o.value, it.value, t, dh.value)


# ---------------------------------------------------------------------------
# Batched ensembles (no counterpart in the reference: it runs one lattice per
# process, main/hmc-nf2.jl)
# ---------------------------------------------------------------------------
class Batch:
    """Independent chains of different lattice sizes and couplings advanced
    together: every CG iteration, Q apply, force and link update is one launch
    over all chains.  Chain c draws its random numbers from (seeds[c], chain c,
    trajectory), as ``HMC`` on a single lattice with ``seed(seeds[c], chain=c)``.
    """

    def __init__(self, sizes, betas, bc=BC_PERIODIC):
        import torch
        n = len(sizes)
        if len(betas) != n:
            raise ValueError("one beta per chain")
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        self.betas = [float(x) for x in betas]
        L1 = (_i * n)(*[a for a, _ in self.sizes])
        L2 = (_i * n)(*[b for _, b in self.sizes])
        h = _p()
        rc = lib.qed2d_batch_create(C.byref(h), n, L1, L2, _arr(self.betas), int(bc),
                                    _p(torch.cuda.current_stream().cuda_stream))
        if rc != QED2D_OK:
            raise QED2dError(rc, "qed2d_batch_create failed")
        self._b = h
        self.n = n
        nchain, sites = _i(), C.c_long()
        lib.qed2d_batch_size(h, C.byref(nchain), C.byref(sites))
        self.sites = sites.value
        self.offsets = []
        for c in range(n):
            o = C.c_long()
            lib.qed2d_batch_offset(h, c, C.byref(o))
            self.offsets.append(o.value)

    def _check(self, rc):
        if rc != QED2D_OK:
            msg = lib.qed2d_batch_last_error(self._b)
            raise (CGNotConverged if rc == QED2D_ENOCONV else QED2dError)(
                rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_b", None) is not None and not _exiting:
            lib.qed2d_batch_destroy(self._b)
        self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reliable_mode(self, on_device=True):
        """Mixed precision: reliable updates requested on the way (True, the
        default) or after a pause (False: each chain reproduces a single-lattice
        solve with set_reliable_mode(prm, False) bit for bit)."""
        self._check(lib.qed2d_batch_set_reliable_mode(self._b, 1 if on_device else 0))

    def set_graphs(self, on=True):
        """Queue each CG chunk as one HIP graph launch (same kernels, same bits)."""
        self._check(lib.qed2d_batch_set_graphs(self._b, 1 if on else 0))

    def arena_spinor(self):
        """A zero arena-shaped spinor (all chains' fields back to back)."""
        import torch
        return torch.zeros(2 * self.sites, dtype=torch.complex128, device="cuda")

    def field(self, t, c):
        """Chain c's (2, L2, L1) view of an arena-shaped field."""
        L1, L2 = self.sizes[c]
        o = 2 * self.offsets[c]
        return t[o:o + 2 * L1 * L2].view(2, L2, L1)

    def get_links(self, c):
        import torch
        L1, L2 = self.sizes[c]
        U = torch.empty((2, L2, L1), dtype=torch.complex128, device="cuda")
        self._check(lib.qed2d_batch_get_links(self._b, c, _ptr(U)))
        return U

    def set_links(self, c, U):
        L1, L2 = self.sizes[c]
        self._check(lib.qed2d_batch_set_links(self._b, c, _ptr(U, None, (2, L2, L1))))

    def set_tiles(self, q64=0, cg64=0, cg32=0, cg64_k2=None, cg32_k2=None):
        """Tile heights of the FP64 apply and the FP64 / FP32 solver kernels (0
        leaves one unchanged); cg64_k2 / cg32_k2 give the second CG kernel's."""
        self._check(lib.qed2d_batch_set_tiles(self._b, int(q64), int(cg64), int(cg32)))
        if cg64_k2 is not None or cg32_k2 is not None:
            cur = self.get_tiles()
            a = cur["cg64"] if cg64_k2 is None else cg64_k2
            b = cur["cg32"] if cg32_k2 is None else cg32_k2
            self._check(lib.qed2d_batch_set_tiles_k2(self._b, int(a), int(b)))

    def get_tiles(self):
        t = (_i * 5)()
        self._check(lib.qed2d_batch_get_tiles(self._b, t))
        return dict(zip(("q64", "cg64", "cg32", "cg64_k2", "cg32_k2"), list(t)))

    def autotune(self, iters=20):
        """Times every tile height on the batch's own lattices and links and
        keeps the fastest (returns the choice, as get_tiles)."""
        t = (_i * 5)()
        self._check(lib.qed2d_batch_autotune(self._b, int(iters), t))
        return self.get_tiles()

    def hot_start(self, seeds):
        s = (_u64 * self.n)(*[int(x) for x in seeds])
        self._check(lib.qed2d_batch_hot_start(self._b, s))

    def CG(self, x, src, am0s, maxiter, tol, precision=PREC_F64):
        """(gamm5Dw)^2 x = src on every chain (x, src arena-shaped); returns the
        iterations per chain."""
        it = (_i * self.n)()
        self._check(lib.qed2d_batch_cg(self._b, _ptr(x), _ptr(src), _arr(am0s), int(maxiter),
                                       float(tol), int(precision), it))
        return list(it)

    def HMC(self, am0s, eps, ns, CGmaxiter, tol, seeds, traj, integrator=None,
            precision=PREC_F64, acc_tol=0.0):
        """One Nf=2 trajectory of every chain; returns a TrajInfo per chain."""
        integ = integrator or Leapfrog()
        n = self.n
        hi, hf = (_d * n)(), (_d * n)()
        acc, its = (_i * n)(), (_i * n)()
        s = (_u64 * n)(*[int(x) for x in seeds])
        self._check(lib.qed2d_batch_hmc_nf2(self._b, _arr(am0s), float(eps), int(ns), integ.code,
                                            int(CGmaxiter), float(tol), int(precision),
                                            float(acc_tol), s, int(traj), hi, hf, acc, its))
        dh = (_d * n)()
        self._check(lib.qed2d_batch_last_dh(self._b, dh))
        return [TrajInfo(hi[c], hf[c], acc[c], its[c], int(traj), dh[c]) for c in range(n)]

    def observables(self):
        """Per-chain (plaquette, topological charge, gauge action) lists."""
        n = self.n
        p, q, a = (_d * n)(), (_d * n)(), (_d * n)()
        self._check(lib.qed2d_batch_observables(self._b, p, q, a))
        return list(p), list(q), list(a)


# ---------------------------------------------------------------------------
# src/QED2dLoadSave.jl
# ---------------------------------------------------------------------------
def save_gauge(U, file, prm):
    """save_gauge (src/QED2dLoadSave.jl:1-10)."""
    _check(prm, lib.qed2d_save_gauge(prm.ctx(), _sp(prm, U), os.fsencode(file)))


def load_gauge(U, file, prm):
    """load_gauge (src/QED2dLoadSave.jl:12-19)."""
    _check(prm, lib.qed2d_load_gauge(prm.ctx(), _sp(prm, U), os.fsencode(file)))


def write_links_text(Uh, file):
    """Host form: Uh is a numpy complex128 array in the (2, L2, L1) layout."""
    a = np.ascontiguousarray(Uh, dtype=np.complex128)
    rc = lib.qed2d_write_links_text(a.ctypes.data_as(_p), a.size, os.fsencode(file))
    if rc != QED2D_OK:
        raise QED2dError(rc, f"cannot write {file}")


def read_links_text(file, L1, L2):
    a = np.zeros((2, L2, L1), np.complex128)
    rc = lib.qed2d_read_links_text(a.ctypes.data_as(_p), a.size, os.fsencode(file))
    if rc != QED2D_OK:
        raise QED2dError(rc, f"cannot read {2 * L1 * L2} links from {file}")
    return a


def config_name(prm, am0, n):
    """File name of main/hmc-nf2.jl:143."""
    return f"config_{prm.iL[0]}_{prm.iL[1]}_b{prm.beta}_m{am0}_n{n}"


# ---------------------------------------------------------------------------
# src/QED2dFields.jl
# ---------------------------------------------------------------------------
def _unfold(Uin, prm, reflect):
    import torch
    out = torch.empty((2, 2 * prm.iL[1], 2 * prm.iL[0]), dtype=torch.complex128,
                      device=Uin.device)
    _check(prm, lib.qed2d_unfold(prm.ctx(), _ptr(out), _sp(prm, Uin), 1 if reflect else 0))
    return out


def unfold_fld(Uin, prm, kprm=None):
    """unfold_fld! (src/QED2dFields.jl:1-10): the field on a 2L1 x 2L2 lattice."""
    return _unfold(Uin, prm, False)


def unfold_reflect(Uin, prm, kprm=None):
    """unfold_reflect! (src/QED2dFields.jl:27-35)."""
    return _unfold(Uin, prm, True)
