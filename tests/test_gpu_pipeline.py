"""run_from_files (corr_amd/pipeline.py) on synthetic PREFIX.* files written in
the reference's layouts, against the same sequence of the CPU oracle.

The two-electron integrals are exactly (mu nu|la si) = sum_Q B_Q[mu nu] B_Q[la si]
and PREFIX.rimp2 holds that B, so RI-MP2 must equal canonical MP2 as well.
Bar: FP64 energies to 1e-10 Eh."""
import sys

import numpy as np
import pytest

from corr_amd import formats as FM
from oracle import ccsd as CC
from oracle import posthf as O
from oracle import scf as SCF

N, NOCC, NAUX = 10, 3, 24


def write_system_files(prefix, guess="core", seed=3):
    rng = np.random.default_rng(seed)
    d = np.concatenate([np.linspace(-2.0, -1.0, NOCC), np.linspace(0.5, 2.5, N - NOCC)])
    h = np.diag(d) + 0.05 * rng.standard_normal((N, N))
    h = 0.5 * (h + h.T)
    B = rng.standard_normal((NAUX, N, N)) * 0.12
    B = 0.5 * (B + B.transpose(0, 2, 1))
    X = np.eye(N) + 0.15 * rng.standard_normal((N, N))
    S = X.T @ X
    H = X.T @ h @ X
    Bao = np.einsum("km,Qkl,ln->Qmn", X, B, X)
    eri = np.einsum("Qmn,Qrs->mnrs", Bao, Bao)
    val, idx, sc = O.sorted_unique_list(eri)
    coord = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.1], [1.7, 0.3, -0.4]])
    charges = np.array([3.0, 1.0, 2.0])
    FM.write_system(prefix + ".sys", FM.SystemInfo(2 * NOCC, 3, N, 0, NAUX, len(val), 0, NAUX, coord, charges,
                                                   charges, np.array([6.0, 1.0, 4.0]), ("SYN", "", "SYN-RI")))
    FM.write_oei(prefix + ".oei", H, np.zeros((N, N)), S, np.zeros((N, N)))
    FM.write_tei(prefix + ".tei", sc, val, idx)
    C0 = np.zeros((N, N)) if guess == "core" else SCF.diag_fmat(H, SCF.calc_s12(S))[0]
    FM.write_wav(prefix + ".ahfw", np.zeros(N), C0)
    FM.write_bpq(prefix + ".rimp2", Bao)
    return H, S, eri, Bao, C0, SCF.calc_ion_rep(coord, charges)


def oracle_run(H, S, eri, Bao, C0, ion_rep):
    T = SCF.calc_s12(S)
    C0, _ = SCF.check_initial_guess(C0, H, T)
    r = SCF.run_scf(H, S, lambda P: SCF.fock_dense(eri, P, H), C0, 2 * NOCC, ion_rep=ion_rep)
    C, F = r["C"], r["F"]
    FMo = O.build_fmo(C, F)
    prec = O.motrans(C, eri)
    mp2 = O.canonical_mp2(prec, FMo, NOCC)
    Bia = O.transform_ri(C, Bao, NOCC)
    e = np.diag(FMo)
    ri = O.rimp2(Bia, e[:NOCC], e[NOCC:])
    cc = CC.ccsd_ur(prec, C, H, 2 * NOCC, ion_rep, maxiter=20)
    return r, mp2, ri, cc


def test_run_from_files_rejects_missing_tei(tmp_path):
    pre = str(tmp_path / "x")
    write_system_files(pre)
    s = FM.read_system(pre + ".sys")
    s.nrofint = 0
    FM.write_system(pre + ".sys", s)
    from corr_amd.pipeline import run_from_files
    with pytest.raises(ValueError, match="libint2"):
        run_from_files(pre, log=None)


@pytest.mark.gpu
@pytest.mark.parametrize("guess", ["core", "given"])
def test_run_from_files_vs_oracle(tmp_path, guess):
    print(f"run_from_files guess={guess}", flush=True)
    sys.stdout.flush()
    from corr_amd.pipeline import run_from_files
    pre = str(tmp_path / "syn")
    H, S, eri, Bao, C0, ion_rep = write_system_files(pre, guess)
    r_ref, mp2_ref, ri_ref, cc_ref = oracle_run(H, S, eri, Bao, C0, ion_rep)
    out = run_from_files(pre, log=None)
    assert abs(out["ion_rep"] - ion_rep) < 1e-14
    assert out["scf"]["core_guess"] == (1 if guess == "core" else 0)
    assert out["scf"]["iterations"] == r_ref["iterations"]
    assert np.abs(out["scf"]["energies"] - r_ref["energies"]).max() < 1e-10
    assert abs(out["scf"]["e_total"] - r_ref["total"]) < 1e-10
    assert abs(out["mp2"]["EMP2"] - mp2_ref["EMP2"]) < 1e-10
    assert abs(out["rimp2"]["EMP2"] - ri_ref["EMP2"]) < 1e-10
    assert abs(out["rimp2"]["EMP2"] - out["mp2"]["EMP2"]) < 1e-10      # exact factorisation
    assert np.abs(np.array(out["ccsd"]["energies"]) - cc_ref["energies"]).max() < 1e-10
    assert abs(out["ccsd"]["e_hf_check"] - r_ref["total"]) < 1e-9
