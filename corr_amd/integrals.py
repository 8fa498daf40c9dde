"""Gaussian integrals on the GPU (libcorr_gpu.so), the counterpart of the
reference's libint2 calls (calculate_libint_oei, ops_rhf.cc:77-103;
calculate_libint_tei, :618-723; calculate_ri, ri.cc:25-100)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._abi import Context, ptr
from .basis import Molecule, ShellArrays


def _index(device) -> int:
    if device is None:
        return torch.cuda.current_device()
    if isinstance(device, int):
        return device
    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError("the integral engine runs on the GPU only")
    return d.index if d.index is not None else torch.cuda.current_device()


def _ctx(device) -> Context:
    return Context.get(_index(device))


def _empty(shape, device):
    return torch.empty(shape, dtype=torch.float64, device=torch.device("cuda", _index(device)))


def one_electron(mol: Molecule, obs: ShellArrays, device=None):
    """(S, T, V) over the orbital basis."""
    n = obs.nbf
    S, T, V = _empty((n, n), device), _empty((n, n), device), _empty((n, n), device)
    q = np.ascontiguousarray(mol.charges, dtype=np.float64)
    X = np.ascontiguousarray(mol.coords, dtype=np.float64).reshape(-1)
    b = obs.struct()
    _ctx(device).call("corr_int1e", C.byref(b), len(q), q.ctypes.data_as(C.c_void_p), X.ctypes.data_as(C.c_void_p),
                      ptr(S), ptr(T), ptr(V))
    return S, T, V


def dipole(obs: ShellArrays, origin=(0.0, 0.0, 0.0), device=None) -> torch.Tensor:
    """AO dipole integrals [3, N, N]: <mu| r_d - origin_d |nu> (psi4's
    MintsHelper.ao_dipole, used by the reference's Boys localisation,
    syshfwplug/pymodule.py:156-160)."""
    D = _empty((3, obs.nbf, obs.nbf), device)
    o = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
    b = obs.struct()
    _ctx(device).call("corr_int1e_dipole", C.byref(b), o.ctypes.data_as(C.c_void_p), ptr(D))
    return D


def metric(aux: ShellArrays, device=None) -> torch.Tensor:
    """(P|Q), naux x naux."""
    J = _empty((aux.nbf, aux.nbf), device)
    b = aux.struct()
    _ctx(device).call("corr_int2c", C.byref(b), ptr(J))
    return J


def three_center(aux: ShellArrays, obs: ShellArrays, device=None) -> torch.Tensor:
    """(P|mu nu) as [naux, N, N]."""
    out = _empty((aux.nbf, obs.nbf, obs.nbf), device)
    ba, bo = aux.struct(), obs.struct()
    _ctx(device).call("corr_int3c", C.byref(ba), C.byref(bo), ptr(out))
    return out


def four_center(obs: ShellArrays, device=None) -> torch.Tensor:
    """Dense (mu nu|la si), N^4."""
    n = obs.nbf
    out = _empty((n, n, n, n), device)
    b = obs.struct()
    _ctx(device).call("corr_int4c", C.byref(b), ptr(out))
    return out
