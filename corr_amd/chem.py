"""Real-chemistry inputs for the post-HF path, computed on the GPU.

A molecule (a seeded water cluster, or one water), its def2-SVP /
def2-SVP-RI integrals (corr_amd/integrals.py), the RI metric (P|Q) and
(P|mu nu), and an RHF (corr_amd/scf.py).  The Fock build is either exact,
over the dense (mu nu|la si) tensor (jk="dense", the reference's own SCF;
N^4 doubles, so up to (H2O)8), or RI-JK from B^Q_mu nu = L^-1 (P|mu nu)
(jk="ri", any size).  The result holds what the reference's main.cc hands to
its post-HF steps: MOs, the AO Fock matrix, H, orbital energies from
FMo = C F C^T (can_mp2.cc:83-91), the metric and (P|mu nu), and with
jk="dense" the dense integral tensor for MOtrans.

The reference's SCF starts from converged psi4 orbitals; started from
scratch, its DIIS gate (mean |FPS - SPF| < 1e-3) never opens for clusters of
8 or more waters (the damped iteration oscillates between two states), so
here DIIS starts at iteration 8 unconditionally and the start is a
superposition of molecular densities: each water's own RI-JK RHF, its
occupied orbitals placed in the cluster basis (guess="sad"), or the core
guess (guess="core").
"""
from __future__ import annotations

import time

import torch

from . import integrals
from .basis import Molecule, ShellArrays, read_g94, water, water_cluster
from .posthf import build_FMo, calculate_ri
from .scf import run_scf, run_scf_ri


def _rhf_inputs(mol, obs, aux, device):
    O, A = ShellArrays(mol, read_g94(obs)), ShellArrays(mol, read_g94(aux))
    S, T, V = integrals.one_electron(mol, O, device)
    H = (T + V).contiguous()
    J = integrals.metric(A, device)
    Pmn = integrals.three_center(A, O, device)
    return O, A, S, H, J, Pmn


def molecular_guess(mol, nmol: int, obs: str, aux: str, device=None) -> torch.Tensor:
    """Orbitals whose occupied rows are each molecule's own converged RHF
    occupied orbitals, placed on that molecule's functions (the atoms of one
    molecule are consecutive, so its functions are a contiguous block)."""
    per = len(mol.symbols) // nmol
    blocks = []
    for m in range(nmol):
        sub = Molecule(mol.symbols[per * m:per * (m + 1)], mol.coords[per * m:per * (m + 1)])
        O, A, S, H, J, Pmn = _rhf_inputs(sub, obs, aux, device)
        BPQ = calculate_ri(J, Pmn, "cholesky", "fp64")
        r = run_scf_ri(H, S, BPQ, torch.zeros_like(S), sub.nelectrons(), e_tol=1e-9, diis_thresh=1e9)
        blocks.append((r["MOs"], sub.nelectrons() // 2, O.nbf))
    N = sum(b[2] for b in blocks)
    C0 = torch.zeros(N, N, dtype=torch.float64, device=blocks[0][0].device)
    row_o, row_v, col = 0, sum(b[1] for b in blocks), 0
    for Cm, no, nb in blocks:
        C0[row_o:row_o + no, col:col + nb] = Cm[:no]
        C0[row_v:row_v + nb - no, col:col + nb] = Cm[no:]
        row_o, row_v, col = row_o + no, row_v + nb - no, col + nb
    return C0


def molecule_rhf(mol: Molecule, nmol: int = 1, obs: str = "def2-svp", aux: str = "def2-svp-ri", device=None,
                 jk: str = "ri", e_tol: float = 1e-10, maxiter: int = 100, guess: str = "sad", log=None) -> dict:
    """RHF of mol (made of nmol equal molecules when guess="sad")."""
    say = log or (lambda *a: None)
    t0 = time.perf_counter()
    O, A, S, H, J, Pmn = _rhf_inputs(mol, obs, aux, device)
    G = integrals.four_center(O, device) if jk == "dense" else None
    torch.cuda.synchronize()
    t_int = time.perf_counter() - t0
    say(f"integrals: N={O.nbf} naux={A.nbf}{' + dense (mu nu|la si)' if G is not None else ''} in {t_int:.2f} s")
    nroe = mol.nelectrons()
    t1 = time.perf_counter()
    C0 = molecular_guess(mol, nmol, obs, aux, device) if guess == "sad" else torch.zeros_like(S)
    kw = dict(ion_rep=mol.nuclear_repulsion(), e_tol=e_tol, maxiter=maxiter, diis_thresh=1e9,
              core_guess_check=(guess != "sad"))
    if jk == "dense":
        scf = run_scf(H, S, G, C0, nroe, **kw)
    else:
        BPQ = calculate_ri(J, Pmn, "cholesky", "fp64")
        scf = run_scf_ri(H, S, BPQ, C0, nroe, **kw)
        del BPQ
    torch.cuda.synchronize()
    t_scf = time.perf_counter() - t1
    say(f"{'RI-JK' if jk == 'ri' else 'exact-integral'} RHF: E = {scf['e_total']:.10f} in {scf['iterations']} "
        f"iterations, {t_scf:.2f} s")
    C, F = scf["MOs"], scf["Fmat"]
    eps = torch.diagonal(build_FMo(C, F)).contiguous()
    nocc = nroe // 2
    return {"mol": mol, "nbf": O.nbf, "naux": A.nbf, "nocc": nocc, "nroe": nroe, "H": H, "S": S, "J": J,
            "Pmn": Pmn, "G": G, "C": C, "F": F, "eps_occ": eps[:nocc].contiguous(),
            "eps_vir": eps[nocc:].contiguous(), "scf": scf, "guess": guess, "jk": jk,
            "seconds_integrals": t_int, "seconds_scf": t_scf}


def water_cluster_rhf(nmol: int, seed: int = 0, obs: str = "def2-svp", aux: str = "def2-svp-ri", device=None,
                      e_tol: float = 1e-10, maxiter: int = 100, guess: str = "sad", jk: str = "ri",
                      log=None) -> dict:
    """RHF of the seeded (H2O)nmol cluster of corr_amd.basis.water_cluster."""
    return molecule_rhf(water_cluster(nmol, seed), nmol, obs, aux, device, jk, e_tol, maxiter, guess, log)


def single_water_rhf(obs: str = "def2-svp", aux: str = "def2-svp-ri", device=None, jk: str = "dense",
                     e_tol: float = 1e-10, log=None) -> dict:
    """One water with O at the origin, O-H 0.958 A, H-O-H 104.5 deg."""
    s, X = water()
    return molecule_rhf(Molecule(s, X), 1, obs, aux, device, jk, e_tol, 100, "core", log)
