"""Orbital localisation on the GPU (libcorr_gpu.so, corr_localize): the
counterpart of localize_PM / localize_boys and calculate_PM / calculate_boys
of syshfwplug/pymodule.py:75-167.

The reference takes a psi4 wavefunction and returns (scipy result, L) with L
the nbf x nocc localised orbitals (columns).  Here the inputs are explicit:
the occupied orbitals C (nocc x nbf, one orbital per row, as run_scf's MOs),
the overlap S and the atom of every basis function (Pipek-Mezey), or the AO
dipole matrices (Boys).  The result dictionary holds the spread before and
after, the number of Jacobi sweeps and whether they converged; L has C's
layout, its order and signs are those the sweeps reach.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._abi import Context, LocOpts, LocResult, ptr

PM, BOYS = 0, 1


def _check(t: torch.Tensor, name: str):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float64 CUDA tensor")


def localize(C_occ: torch.Tensor, kind: str, S: torch.Tensor | None = None, bf_atom=None,
             D: torch.Tensor | None = None, max_sweeps: int = 100, tol: float = 1e-14):
    """kind = "pm" (needs S and bf_atom) or "boys" (needs D, [3, nbf, nbf])."""
    _check(C_occ, "C_occ")
    n, nbf = C_occ.shape
    kinds = {"pm": PM, "boys": BOYS}
    if kind not in kinds:
        raise ValueError("kind must be 'pm' or 'boys'")
    if kind == "pm":
        if S is None or bf_atom is None:
            raise ValueError("Pipek-Mezey needs S and bf_atom")
        _check(S, "S")
        atoms = np.ascontiguousarray(bf_atom, dtype=np.int32)
        if atoms.shape != (nbf,):
            raise ValueError("bf_atom needs one entry per basis function")
        natom = int(atoms.max()) + 1
        M, ap = S, atoms.ctypes.data_as(C.c_void_p)
    else:
        if D is None:
            raise ValueError("Boys needs the dipole matrices D")
        _check(D, "D")
        if tuple(D.shape) != (3, nbf, nbf):
            raise ValueError("D must be [3, nbf, nbf]")
        natom, M, ap, atoms = 0, D, None, None
    L = torch.empty_like(C_occ)
    res = LocResult()
    opts = LocOpts(max_sweeps, tol)
    ctx = Context.get(C_occ.device.index if C_occ.device.index is not None else torch.cuda.current_device())
    ctx.call("corr_localize", nbf, n, ptr(C_occ), kinds[kind], ptr(M), ap, natom, C.byref(opts), ptr(L),
             C.byref(res))
    return {"spread_initial": res.spread_initial, "spread": res.spread, "sweeps": res.sweeps,
            "converged": bool(res.converged)}, L


def localize_PM(C_occ: torch.Tensor, S: torch.Tensor, bf_atom, **kw):
    """localize_PM (pymodule.py:75-97): (result, L)."""
    return localize(C_occ, "pm", S=S, bf_atom=bf_atom, **kw)


def localize_boys(C_occ: torch.Tensor, D: torch.Tensor, **kw):
    """localize_boys (pymodule.py:103-124): (result, L)."""
    return localize(C_occ, "boys", D=D, **kw)
