"""PyTorch-ROCm front-end of the spin-orbital CCSD (src/POSTHF/ccsd.cc,
src/POSTHF/ccsd_intermediates.cc).  All work runs in libcorr_gpu.so."""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from ._abi import CCSDOpts, CCSDResult, Context, ptr
from .posthf import _ctx, _dev, _prec

MODES = {"reference": _abi.CCSD_REFERENCE, "diis": _abi.CCSD_DIIS}


def ints_so(prec_ints: torch.Tensor) -> torch.Tensor:
    """<pq||rs> over 2N spin orbitals (get_integral / calc_ints_so, ccsd.cc:6-9, 509-535)."""
    _dev(prec_ints)
    N = prec_ints.shape[0]
    out = torch.empty((2 * N,) * 4, dtype=torch.float64, device=prec_ints.device)
    _ctx(prec_ints).call("corr_ccsd_ints_so", N, ptr(prec_ints), ptr(out))
    return out


def ccsd_update(V: torch.Tensor, f: torch.Tensor, T1: torch.Tensor, T2: torch.Tensor,
                prec: str | int = "fp64") -> tuple[torch.Tensor, torch.Tensor]:
    """One Jacobi update (T1, T2) -> (T1n, T2n): taus, intermediates, T1 and T2
    residuals divided by the denominators (ccsd.cc:221-467)."""
    _dev(V, f, T1, T2)
    o, v = T1.shape
    T1n, T2n = torch.empty_like(T1), torch.empty_like(T2)
    _ctx(V).call("corr_ccsd_update", o, v, ptr(V), ptr(f), ptr(T1), ptr(T2), ptr(T1n), ptr(T2n), _prec(prec))
    return T1n, T2n


def ccsd_energy(V: torch.Tensor, f: torch.Tensor, T1: torch.Tensor, T2: torch.Tensor) -> float:
    """E(CCSD) of ccsd_energy (ccsd.cc:147-219)."""
    _dev(V, f, T1, T2)
    o, v = T1.shape
    e = C.c_double(0.0)
    _ctx(V).call("corr_ccsd_energy", o, v, ptr(V), ptr(f), ptr(T1), ptr(T2), C.byref(e))
    return e.value


def ccsd_ur(prec_ints: torch.Tensor, MOs: torch.Tensor, Hmat: torch.Tensor, nroe: int, ion_rep: float,
            maxiter: int = 20, mode: str = "reference", diis_size: int = 8, prec: str | int = "fp64",
            e_tol: float = 1e-8, r_tol: float = 1e-6) -> dict:
    """ccsd_ur (ccsd.cc:538-580).  mode='reference' runs maxiter Jacobi updates
    exactly like the reference; mode='diis' adds DIIS and stops at |dE| < e_tol
    and amplitude-change norm < r_tol."""
    _dev(prec_ints, MOs, Hmat)
    N = MOs.shape[0]
    o, v = nroe, 2 * N - nroe
    T1 = torch.empty(o, v, dtype=torch.float64, device=MOs.device)
    T2 = torch.empty(o, o, v, v, dtype=torch.float64, device=MOs.device)
    energies = (C.c_double * maxiter)()
    opts = CCSDOpts(MODES[mode], maxiter, diis_size, _prec(prec), e_tol, r_tol)
    res = CCSDResult()
    _ctx(MOs).call("corr_ccsd", N, nroe, ptr(prec_ints), ptr(MOs), ptr(Hmat), ion_rep, C.byref(opts),
                   energies, ptr(T1), ptr(T2), C.byref(res))
    return {"energies": list(energies)[: res.iterations], "emp2": res.e_mp2, "e_hf_check": res.e_hf_check,
            "e_ccsd": res.e_ccsd, "d_e": res.d_e, "r_norm": res.r_norm, "iterations": res.iterations,
            "converged": bool(res.converged), "T1": T1, "T2": T2}
