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


if __name__ == "__main__":
    rimp2_case()
    fourindex_case()
    ccsd_case()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
