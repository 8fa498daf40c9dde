"""PyTorch-ROCm front-end of the post-HF path (names follow src/POSTHF).

Tensors are FP64 on a gfx950 device; every function enqueues libcorr_gpu.so
kernels on torch's current stream.  There is no CPU fallback: a CPU tensor or
a missing library raises.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from ._abi import Context, MP2Energy, ptr


def _prec(prec: str | int) -> int:
    if isinstance(prec, int):
        return prec
    try:
        return _abi.PRECISIONS[prec]
    except KeyError:
        raise ValueError(f"unknown precision {prec!r}; use one of {list(_abi.PRECISIONS)}") from None


def _gemm_prec(prec: str | int) -> int:
    if isinstance(prec, int):
        return prec
    try:
        return _abi.GEMM_PRECISIONS[prec]
    except KeyError:
        raise ValueError(f"unknown GEMM precision {prec!r}; use one of {list(_abi.GEMM_PRECISIONS)}") from None


def _dev(*ts: torch.Tensor) -> torch.Tensor:
    for t in ts:
        if not isinstance(t, torch.Tensor):
            raise TypeError("expected torch tensors")
        if t.device.type != "cuda":
            raise ValueError("corr_amd runs on the GPU only: move inputs to a cuda (HIP) device")
        if t.dtype != torch.float64:
            raise ValueError(f"expected float64 tensors, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError("expected contiguous tensors")
    return ts[0]


def _ctx(t: torch.Tensor) -> Context:
    return Context.get(t.device.index if t.device.index is not None else torch.cuda.current_device())


def dgemm(A: torch.Tensor, B: torch.Tensor, transa: bool = False, transb: bool = False,
          alpha: float = 1.0, prec: str | int = "fp64") -> torch.Tensor:
    """C = alpha op(A) op(B) on the MFMA GEMM (2-D, row-major).  prec "int8x5":
    FP64 emulated on the int8 matrix cores."""
    _dev(A, B)
    M = A.shape[1] if transa else A.shape[0]
    K = A.shape[0] if transa else A.shape[1]
    N = B.shape[0] if transb else B.shape[1]
    Kb = B.shape[1] if transb else B.shape[0]
    if K != Kb:
        raise ValueError("inner dimensions differ")
    Cm = torch.empty(M, N, dtype=torch.float64, device=A.device)
    _ctx(A).call("corr_dgemm", int(transa), int(transb), M, N, K, alpha, ptr(A), A.shape[1], 0,
                 ptr(B), B.shape[1], 0, 0.0, ptr(Cm), N, 0, 1, _gemm_prec(prec))
    return Cm


def build_FMo(MOs: torch.Tensor, Fmat: torch.Tensor) -> torch.Tensor:
    """FMo = C F C^T (motrans.cc:8-32)."""
    _dev(MOs, Fmat)
    N = MOs.shape[0]
    FMo = torch.empty(N, N, dtype=torch.float64, device=MOs.device)
    _ctx(MOs).call("corr_build_fmo", N, ptr(MOs), ptr(Fmat), ptr(FMo))
    return FMo


def ao_unfold(nroao: int, intval: torch.Tensor, intnums: torch.Tensor, sortcount) -> torch.Tensor:
    """Dense N^4 AO tensor from the sorted unique list (trans_I semantics)."""
    _dev(intval)
    if intnums.dtype not in (torch.int16, torch.uint16) or intnums.device != intval.device:
        raise ValueError("intnums must be a 16-bit integer tensor on the same device")
    sc = (C.c_int64 * 4)(*[int(x) for x in sortcount])
    dense = torch.empty((nroao,) * 4, dtype=torch.float64, device=intval.device)
    _ctx(intval).call("corr_ao_unfold", nroao, intval.numel(), sc, ptr(intval), ptr(intnums), ptr(dense))
    return dense


def MOtrans(MOs: torch.Tensor, dense_ao: torch.Tensor, prec: str | int = "fp64") -> torch.Tensor:
    """prec_ints (ij|kl) from the dense AO tensor (MOtrans, motrans.cc:35-211)."""
    _dev(MOs, dense_ao)
    N = MOs.shape[0]
    work = torch.empty_like(dense_ao)
    out = torch.empty_like(dense_ao)
    _ctx(MOs).call("corr_motrans", N, ptr(MOs), ptr(dense_ao), ptr(work), ptr(out), _prec(prec))
    return out


def run_canonical_mp2(prec_ints: torch.Tensor, FMo: torch.Tensor, nocc: int) -> dict:
    """EMP2 from prec_ints (can_mp2.cc:11-49)."""
    _dev(prec_ints, FMo)
    e = MP2Energy()
    _ctx(FMo).call("corr_canonical_mp2", FMo.shape[0], nocc, ptr(prec_ints), ptr(FMo), C.byref(e))
    return e.as_dict()


def run_non_canonical_mp2(prec_ints: torch.Tensor, FMo: torch.Tensor, nocc: int, maxiter: int = 20,
                          return_T: bool = False) -> dict:
    """Semi-canonical MP2 with the off-diagonal occupied Fock block
    (run_non_canonical_mp2, loc_mp2.cc:7-92): per-iteration "E(sem-loc)" and
    "Rsum"; optionally the final amplitudes T[i][j][a][b]."""
    import numpy as np
    _dev(prec_ints, FMo)
    N = FMo.shape[0]
    v = N - nocc
    E = np.zeros(max(maxiter, 1))
    R = np.zeros(max(maxiter, 1))
    T = torch.empty(nocc, nocc, v, v, dtype=torch.float64, device=FMo.device) if return_T else None
    _ctx(FMo).call("corr_non_canonical_mp2", N, nocc, ptr(prec_ints), ptr(FMo), maxiter,
                   E.ctypes.data_as(C.c_void_p), R.ctypes.data_as(C.c_void_p), ptr(T))
    out = {"energies": E[:maxiter].copy(), "rsum": R[:maxiter].copy()}
    if return_T:
        out["T"] = T
    return out


def ri_metric(J: torch.Tensor, mode: str = "eigen") -> tuple[torch.Tensor, int]:
    """(M, kept) with M^T M = J^-1 (ri.cc:50-69 for mode='eigen')."""
    _dev(J)
    naux = J.shape[0]
    M = torch.empty_like(J)
    kept = C.c_int64(0)
    _ctx(J).call("corr_ri_metric", naux, ptr(J), _abi.METRICS[mode], ptr(M), C.byref(kept))
    return M, kept.value


def calculate_ri(J: torch.Tensor, Pmn: torch.Tensor, mode: str = "eigen",
                 prec: str | int = "fp64") -> torch.Tensor:
    """BPQ = J^-1/2 (P|mu nu) from the metric and 3-index integrals
    (arithmetic of calculate_ri, ri.cc:50-69, 102-103)."""
    _dev(J, Pmn)
    M, _ = ri_metric(J, mode)
    naux, N = Pmn.shape[0], Pmn.shape[1]
    BPQ = torch.empty_like(Pmn)
    _ctx(J).call("corr_ri_bpq", naux, N, ptr(M), ptr(Pmn), ptr(BPQ), _prec(prec))
    return BPQ


def transform_ri(MOs: torch.Tensor, BPQ: torch.Tensor, nocc: int, prec: str | int = "fp64") -> torch.Tensor:
    """Bia[Q,i,a] (transform_ri, ri.cc:147-172)."""
    _dev(MOs, BPQ)
    naux, N = BPQ.shape[0], BPQ.shape[1]
    Bia = torch.empty(naux, nocc, N - nocc, dtype=torch.float64, device=BPQ.device)
    _ctx(BPQ).call("corr_transform_ri", naux, N, nocc, ptr(MOs), ptr(BPQ), ptr(Bia), _prec(prec))
    return Bia


def ri_form_bia(MOs: torch.Tensor, Pmn: torch.Tensor, M: torch.Tensor, nocc: int,
                mode: str = "cholesky", prec: str | int = "fp64") -> torch.Tensor:
    """Bia = M (P|ia): AO->MO first, metric last (1.97 vs 7.8 TFLOP at (H2O)32)."""
    _dev(MOs, Pmn, M)
    naux, N = Pmn.shape[0], Pmn.shape[1]
    Bia = torch.empty(naux, nocc, N - nocc, dtype=torch.float64, device=Pmn.device)
    _ctx(Pmn).call("corr_ri_form_bia", naux, N, nocc, ptr(MOs), ptr(Pmn), ptr(M), _abi.METRICS[mode],
                   ptr(Bia), _prec(prec))
    return Bia


def ri_bia_from_j(MOs: torch.Tensor, Pmn: torch.Tensor, J: torch.Tensor, nocc: int,
                  mode: str = "cholesky", prec: str | int = "fp64") -> tuple[torch.Tensor, int]:
    """(Bia, kept) from J and (P|mu nu): the metric factorisation runs on a
    second stream, overlapped with the AO->MO GEMMs (calculate_ri +
    transform_ri arithmetic, ri.cc:50-69, 102-103, 147-172)."""
    _dev(MOs, Pmn, J)
    naux, N = Pmn.shape[0], Pmn.shape[1]
    Bia = torch.empty(naux, nocc, N - nocc, dtype=torch.float64, device=Pmn.device)
    kept = C.c_int64(0)
    _ctx(Pmn).call("corr_ri_bia_from_j", naux, N, nocc, ptr(MOs), ptr(Pmn), ptr(J), _abi.METRICS[mode], None,
                   ptr(Bia), _prec(prec), C.byref(kept))
    return Bia, kept.value


def ri_mp2_energy_from_j(MOs: torch.Tensor, Pmn: torch.Tensor, J: torch.Tensor, nocc: int,
                         eps_occ: torch.Tensor, eps_vir: torch.Tensor, mode: str = "cholesky",
                         prec: str | int = "fp64") -> tuple[dict, int]:
    """(energies, kept): B^Q_ia formation from J and (P|mu nu) followed by the
    RI-MP2 energy, without handing B back (calculate_ri + transform_ri,
    ri.cc:50-69, 102-103, 147-172, then run_canonical_mp2_ri,
    can_mp2.cc:53-132).  Same numbers as ri_bia_from_j(..., prec) followed by
    rimp2_energy(Bia, ..., prec); with mode "cholesky" and prec "bf16x3" the
    metric product writes B straight into the energy kernel's BF16 hi/lo
    panels instead of an FP64 B."""
    _dev(MOs, Pmn, J, eps_occ, eps_vir)
    naux, N = Pmn.shape[0], Pmn.shape[1]
    if eps_occ.numel() != nocc or eps_vir.numel() != N - nocc:
        raise ValueError("orbital energy lengths do not match nocc and the basis size")
    e = MP2Energy()
    kept = C.c_int64(0)
    _ctx(Pmn).call("corr_ri_mp2_energy_from_j", naux, N, nocc, ptr(MOs), ptr(Pmn), ptr(J), _abi.METRICS[mode],
                   ptr(eps_occ), ptr(eps_vir), _prec(prec), C.byref(e), C.byref(kept))
    return e.as_dict(), kept.value


def rimp2_energy(Bia: torch.Tensor, eps_occ: torch.Tensor, eps_vir: torch.Tensor,
                 prec: str | int = "fp64") -> dict:
    """RI-MP2 energies (run_canonical_mp2_ri, can_mp2.cc:53-132)."""
    _dev(Bia, eps_occ, eps_vir)
    naux, o, v = Bia.shape
    if eps_occ.numel() != o or eps_vir.numel() != v:
        raise ValueError("orbital energy lengths do not match Bia")
    e = MP2Energy()
    _ctx(Bia).call("corr_rimp2_energy", naux, o, v, ptr(Bia), ptr(eps_occ), ptr(eps_vir), _prec(prec),
                   C.byref(e))
    return e.as_dict()


def run_canonical_mp2_ri(Bia: torch.Tensor, FMo: torch.Tensor, nocc: int, prec: str | int = "fp64") -> dict:
    """Reference-shaped entry: orbital energies from the FMo diagonal
    (can_mp2.cc:83-91)."""
    e = torch.diagonal(FMo).contiguous()
    return rimp2_energy(Bia, e[:nocc].contiguous(), e[nocc:].contiguous(), prec)


def rimp2_gradient(Bia: torch.Tensor, eps_occ: torch.Tensor, eps_vir: torch.Tensor, c_os: float = 1.0,
                   c_ss: float = 1.0, prec: str | int = "fp64"):
    """(E, dE/dBia, dE/deps_occ, dE/deps_vir) for E = c_os EMP2_OS + c_ss EMP2_SS."""
    _dev(Bia, eps_occ, eps_vir)
    naux, o, v = Bia.shape
    dB = torch.empty_like(Bia)
    deo, dev = torch.empty_like(eps_occ), torch.empty_like(eps_vir)
    e = MP2Energy()
    _ctx(Bia).call("corr_rimp2_gradient", naux, o, v, ptr(Bia), ptr(eps_occ), ptr(eps_vir), c_os, c_ss,
                   _prec(prec), ptr(dB), ptr(deo), ptr(dev), C.byref(e))
    return c_os * e.e_os + c_ss * e.e_ss, dB, deo, dev


class RIMP2Energy(torch.autograd.Function):
    """E = c_os EMP2_OS + c_ss EMP2_SS as a differentiable function of
    (Bia, eps_occ, eps_vir).  Forward: fused pair-block energy kernel.
    Backward: pair-block recomputation + gradient kernels (no o^2 v^2
    amplitudes kept between forward and backward)."""

    @staticmethod
    def forward(ctx, Bia, eps_occ, eps_vir, prec="fp64", c_os=1.0, c_ss=1.0, grad_prec=None):
        e = rimp2_energy(Bia, eps_occ, eps_vir, prec)
        ctx.save_for_backward(Bia, eps_occ, eps_vir)
        ctx.c = (c_os, c_ss, grad_prec if grad_prec is not None else prec)
        return torch.tensor(c_os * e["EMP2_OS"] + c_ss * e["EMP2_SS"], dtype=torch.float64, device=Bia.device)

    @staticmethod
    def backward(ctx, gE):
        Bia, eo, ev = ctx.saved_tensors
        c_os, c_ss, prec = ctx.c
        _, dB, deo, dev = rimp2_gradient(Bia, eo, ev, c_os, c_ss, prec)
        g = gE.to(torch.float64)
        return dB * g, deo * g, dev * g, None, None, None, None


def rimp2(Bia, eps_occ, eps_vir, prec="fp64", c_os=1.0, c_ss=1.0, grad_prec=None):
    """Differentiable RI-MP2 energy (see RIMP2Energy)."""
    return RIMP2Energy.apply(Bia, eps_occ, eps_vir, prec, c_os, c_ss, grad_prec)
