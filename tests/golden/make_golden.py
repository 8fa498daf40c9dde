"""Writes the golden fixtures of this directory from the CPU oracle.

The reference ships no logs, fixtures or known-answer tests, so these are the
oracle's FP64 outputs on seeded inputs.  The inputs are stored too, so the
fixtures do not depend on a random-number generator.  Together with the pins
of tests/test_oracle*.py, they hold the oracle and the GPU path to fixed
numbers.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ccsd as CC           # noqa: E402
from oracle import posthf as O          # noqa: E402
from tests import _synth as S           # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def rimp2_case():
    """RI-MP2 energies and gradient (can_mp2.cc:53-132) for naux=76, o=5, v=19."""
    B = S.bia(76, 5, 19, seed=2024)
    eo, ev = S.orbital_energies(5, 19, seed=2024)
    e = O.rimp2(B, eo, ev)
    E, dB, deo, dev = O.rimp2_gradient(B, eo, ev, 1.0, 1.0)
    np.savez(os.path.join(HERE, "rimp2_naux76_o5_v19.npz"), Bia=B, eps_occ=eo, eps_vir=ev,
             e_os=e["EMP2_OS"], e_ss=e["EMP2_SS"], e_mp2=e["EMP2"], e_scs=e["EMP2_SCS"],
             grad_E=E, grad_B=dB, grad_eo=deo, grad_ev=dev)


def fourindex_case():
    """Sorted unique list -> dense AO -> (ij|kl) -> canonical MP2 (motrans.cc,
    can_mp2.cc:11-49) for N=6, 2 occupied."""
    N, o = 6, 2
    eri = S.eri_like(N, seed=606)
    val, idx, sc = O.sorted_unique_list(eri)
    C = S.orthonormal(N, seed=607)
    dense = O.unfold_sorted(N, val, idx, sc)
    prec = O.motrans(C, dense)
    eo, ev = S.orbital_energies(o, N - o, seed=608)
    FMo = np.diag(np.concatenate([eo, ev]))
    e = O.canonical_mp2(prec, FMo, o)
    np.savez(os.path.join(HERE, "fourindex_n6.npz"), intval=val, intnums=idx.astype(np.int32),
             sortcount=sc, C=C, prec_ints=prec, FMo=FMo, e_os=e["EMP2_OS"], e_ss=e["EMP2_SS"],
             e_mp2=e["EMP2"])


def ccsd_case():
    """ccsd_ur (ccsd.cc:538-580): 20 fixed iterations on a 6-orbital model
    Hamiltonian with 4 electrons (RHF orbitals in an orthonormal basis)."""
    N, nocc_sp = 6, 2
    h_ao, eri = S.model_hamiltonian(N, seed=61)
    C, eps, F = S.rhf_orthonormal(h_ao, eri, nocc_sp)
    prec = O.motrans(C, eri)
    r = CC.ccsd_ur(prec, C, h_ao, 2 * nocc_sp, 0.75, maxiter=20)
    np.savez(os.path.join(HERE, "ccsd_n6_e4.npz"), prec_ints=prec, C=C, H=h_ao, nroe=2 * nocc_sp, ion_rep=0.75,
             energies=np.array(r["energies"]), emp2=r["emp2"], e_hf_check=r["e_hf_check"], T1=r["T1"], T2=r["T2"])


def water_case():
    """RI-MP2 of one water (O-H 0.958 A, H-O-H 104.5 deg) in def2-SVP /
    def2-SVP-RI: the oracle's integrals, RHF to |dE| <= 1e-12 (dense Fock),
    J^-1/2 with the reference's eigenvalue cutoff (ri.cc:50-69), B by
    transform_ri (ri.cc:147-172), energies by run_canonical_mp2_ri
    (can_mp2.cc:53-132).  A real-chemistry B: the reduced-precision paths
    are checked on it in micro-hartree."""
    from oracle import integrals as I
    from oracle import scf as SCF
    r, th = 0.958 * I.ANGSTROM_TO_BOHR, np.deg2rad(104.5)
    X = np.array([[0, 0, 0], [r, 0, 0], [r * np.cos(th), r * np.sin(th), 0]])
    atoms, Z = list(zip(["O", "H", "H"], X)), [8.0, 1.0, 1.0]
    bas = os.path.join(ROOT, "corr_amd", "basis")
    sh = I.build_shells(atoms, I.parse_g94(os.path.join(bas, "def2-svp.g94")))
    ax = I.build_shells(atoms, I.parse_g94(os.path.join(bas, "def2-svp-ri.g94")))
    S, T, V, G = I.overlap(sh), I.kinetic(sh), I.nuclear(sh, Z, list(X)), I.coulomb4(sh)
    H = T + V
    res = SCF.run_scf(H, S, lambda P: SCF.fock_dense(G, P, H), SCF.diag_fmat(H, SCF.calc_s12(S))[0], 10,
                      ion_rep=I.nuclear_repulsion(Z, list(X)), e_tol=1e-12, maxiter=200)
    M, kept = O.ri_metric_eigen(I.coulomb2(ax))
    B = O.transform_ri(res["C"], O.ri_bpq(M, I.coulomb3(ax, sh)), 5)
    eps = np.diag(O.build_fmo(res["C"], res["F"]))
    e = O.rimp2(B, eps[:5], eps[5:])
    np.savez(os.path.join(HERE, "rimp2_water_def2svp.npz"), Bia=B, eps_occ=eps[:5], eps_vir=eps[5:],
             e_rhf=res["total"], naux_kept=kept, e_os=e["EMP2_OS"], e_ss=e["EMP2_SS"], e_mp2=e["EMP2"],
             e_scs=e["EMP2_SCS"])


if __name__ == "__main__":
    water_case()
    rimp2_case()
    fourindex_case()
    ccsd_case()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
