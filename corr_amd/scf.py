"""RHF input side on the GPU: Fock build, density and the damped RHF + DIIS
driver (names follow src/SCF).  FP64 tensors on a gfx950 device; no CPU
fallback."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._abi import Context, SCFOpts, SCFResult, ptr
from .posthf import _ctx, _dev


def build_Fmat(dense_ao: torch.Tensor, Pmat: torch.Tensor, Hmat: torch.Tensor) -> torch.Tensor:
    """F = H + J(P) - 1/2 K(P) from the dense AO tensor of ao_unfold
    (build_Fmat, src/SCF/ops_rhf.cc:549-585)."""
    _dev(dense_ao, Pmat, Hmat)
    N = Hmat.shape[0]
    if dense_ao.numel() != N ** 4 or Pmat.shape != (N, N) or Hmat.shape != (N, N):
        raise ValueError("build_Fmat: shapes must be (N,N,N,N), (N,N), (N,N)")
    F = torch.empty_like(Hmat)
    _ctx(Hmat).call("corr_build_fmat", N, ptr(dense_ao), ptr(Pmat), ptr(Hmat), ptr(F))
    return F


def build_Fmat_list(nroao: int, intval: torch.Tensor, intnums: torch.Tensor, sortcount, Pmat: torch.Tensor,
                    Hmat: torch.Tensor) -> torch.Tensor:
    """The reference's argument form: the sorted unique list is unfolded into
    context scratch on every call (ops_rhf.cc:549)."""
    _dev(intval, Pmat, Hmat)
    if intnums.dtype not in (torch.int16, torch.uint16) or intnums.device != intval.device:
        raise ValueError("intnums must be a 16-bit integer tensor on the same device")
    sc = (C.c_int64 * 4)(*[int(x) for x in sortcount])
    F = torch.empty_like(Hmat)
    _ctx(Hmat).call("corr_build_fmat_list", nroao, intval.numel(), sc, ptr(intval), ptr(intnums), ptr(Pmat),
                    ptr(Hmat), ptr(F))
    return F


def build_Pmat(MOs: torch.Tensor, nroe: int, damp: float = 0.0, Pmat: torch.Tensor | None = None) -> torch.Tensor:
    """P = 2(1-damp) C_occ^T C_occ + damp P (build_Pmat, src/SCF/scf.cc:44-61);
    updates Pmat in place when given."""
    _dev(MOs)
    N = MOs.shape[0]
    if Pmat is None:
        if damp != 0.0:
            raise ValueError("damping needs the previous density")
        Pmat = torch.empty_like(MOs)
    else:
        _dev(Pmat)
    _ctx(MOs).call("corr_build_pmat", N, nroe, ptr(MOs), float(damp), ptr(Pmat))
    return Pmat


def _run(fn, args, Hmat, Smat, MOs, nroe, ion_rep, maxiter, diis_size, diis_thresh, damp, e_tol, core_guess_check):
    N = Hmat.shape[0]
    C_out = MOs.clone()
    eps = torch.empty(N, dtype=torch.float64, device=Hmat.device)
    P = torch.empty_like(Hmat)
    F = torch.empty_like(Hmat)
    opts = SCFOpts(maxiter, diis_size, int(core_guess_check), 0, diis_thresh, damp, e_tol)
    energies = np.zeros(max(maxiter, 1))
    errors = np.zeros(max(maxiter, 1))
    res = SCFResult()
    _ctx(Hmat).call(fn, N, nroe, ptr(Hmat), ptr(Smat), *args, ptr(C_out), ptr(eps), ptr(P), ptr(F), float(ion_rep),
                    C.byref(opts), energies.ctypes.data_as(C.c_void_p), errors.ctypes.data_as(C.c_void_p),
                    C.byref(res))
    out = {k: getattr(res, k) for k, _ in SCFResult._fields_}
    out.update(MOs=C_out, MOens=eps, Pmat=P, Fmat=F, energies=energies[:res.iterations].copy(),
               errors=errors[:res.iterations].copy())
    return out


def run_scf(Hmat: torch.Tensor, Smat: torch.Tensor, dense_ao: torch.Tensor, MOs: torch.Tensor, nroe: int,
            ion_rep: float = 0.0, maxiter: int = 100, diis_size: int = 8, diis_thresh: float = 1e-3,
            damp: float = 0.5, e_tol: float = 1e-8, core_guess_check: bool = True) -> dict:
    """Damped RHF + DIIS (run_scf, src/SCF/scf.cc:145-320) with calc_S12 and
    check_initial_guess before it.  MOs (row = MO) are the starting orbitals;
    the result holds the final MOs, MOens, Pmat and Fmat (rebuilt after the
    loop, as the reference leaves them for build_FMo) and the energy lines."""
    _dev(Hmat, Smat, dense_ao, MOs)
    N = Hmat.shape[0]
    if dense_ao.numel() != N ** 4 or Smat.shape != (N, N) or MOs.shape != (N, N):
        raise ValueError("run_scf: shapes must be H, S, MOs (N,N) and dense (N,N,N,N)")
    return _run("corr_run_scf", (ptr(dense_ao),), Hmat, Smat, MOs, nroe, ion_rep, maxiter, diis_size, diis_thresh,
                damp, e_tol, core_guess_check)


def build_Fmat_ri(BPQ: torch.Tensor, Pmat: torch.Tensor, Hmat: torch.Tensor) -> torch.Tensor:
    """RI-JK Fock matrix H + J - 1/2 K from B^Q_mu nu (no reference counterpart)."""
    _dev(BPQ, Pmat, Hmat)
    naux, N = BPQ.shape[0], Hmat.shape[0]
    if BPQ.shape != (naux, N, N):
        raise ValueError("BPQ must be [naux, N, N]")
    F = torch.empty_like(Hmat)
    _ctx(Hmat).call("corr_build_fmat_ri", N, naux, ptr(BPQ), ptr(Pmat), ptr(Hmat), ptr(F))
    return F


def run_scf_ri(Hmat: torch.Tensor, Smat: torch.Tensor, BPQ: torch.Tensor, MOs: torch.Tensor, nroe: int,
               ion_rep: float = 0.0, maxiter: int = 100, diis_size: int = 8, diis_thresh: float = 1e-3,
               damp: float = 0.5, e_tol: float = 1e-8, core_guess_check: bool = True) -> dict:
    """run_scf with the RI-JK Fock build (for clusters whose full integral
    tensor does not fit)."""
    _dev(Hmat, Smat, BPQ, MOs)
    naux, N = BPQ.shape[0], Hmat.shape[0]
    if BPQ.shape != (naux, N, N) or Smat.shape != (N, N) or MOs.shape != (N, N):
        raise ValueError("run_scf_ri: shapes must be H, S, MOs (N,N) and BPQ (naux,N,N)")
    return _run("corr_run_scf_ri", (naux, ptr(BPQ)), Hmat, Smat, MOs, nroe, ion_rep, maxiter, diis_size,
                diis_thresh, damp, e_tol, core_guess_check)


__all__ = ["build_Fmat", "build_Fmat_list", "build_Fmat_ri", "build_Pmat", "run_scf", "run_scf_ri", "Context"]
