"""Pins the CCSD oracle: guess energy == spatial MP2, two-electron CCSD == FCI,
antisymmetry of T2, vanishing converged residual, and the reference's
20-iteration driver shape."""
import numpy as np
import pytest

from oracle import ccsd as CC
from oracle import fci
from oracle import posthf as O
from tests import _synth as S


def _system(N, nocc_sp, seed):
    h_ao, eri = S.model_hamiltonian(N, seed=seed)
    C, eps, F = S.rhf_orthonormal(h_ao, eri, nocc_sp)
    prec = O.motrans(C, eri)
    return h_ao, eri, C, F, prec


def test_spin_orbital_mp2_equals_spatial_mp2():
    N, nocc_sp = 6, 2
    h_ao, eri, C, F, prec = _system(N, nocc_sp, seed=3)
    V = CC.ints_so(prec)
    h, f = CC.form_fock(C, h_ao, V, 2 * nocc_sp)
    B = CC.Blocks(V, f, 2 * nocc_sp)
    _, T2, emp2 = CC.guess(B)
    FMo = O.build_fmo(C, F)
    assert np.allclose(np.diag(f)[::2], np.diag(FMo), atol=1e-10)   # same orbital energies
    spatial = O.canonical_mp2(prec, FMo, nocc_sp)["EMP2"]
    assert np.isclose(emp2, spatial, rtol=1e-10, atol=1e-13)
    # iteration-0 CCSD energy with T1 = 0 is the MP2 energy
    assert np.isclose(CC.energy(B, np.zeros((B.o, B.v)), T2), emp2, rtol=1e-13)
    assert np.allclose(T2, -T2.transpose(1, 0, 2, 3)) and np.allclose(T2, -T2.transpose(0, 1, 3, 2))


def test_ints_so_antisymmetry_and_spin_blocks():
    prec = S.eri_like(4, seed=2)
    V = CC.ints_so(prec)
    assert np.allclose(V, -V.transpose(1, 0, 2, 3)) and np.allclose(V, -V.transpose(0, 1, 3, 2))
    assert np.allclose(V, V.transpose(2, 3, 0, 1))
    # <0a 1b || 0a 1b> (alpha, beta) = (00|11): no exchange between opposite spins
    assert np.isclose(V[0, 3, 0, 3], prec[0, 0, 1, 1])
    # same spin: <0a 2a||0a 2a> = (00|11) - (01|10)
    assert np.isclose(V[0, 2, 0, 2], prec[0, 0, 1, 1] - prec[0, 1, 1, 0])


@pytest.mark.parametrize("N,seed", [(3, 5), (5, 6)])
def test_two_electron_ccsd_equals_fci(N, seed):
    h_ao, eri, C, F, prec = _system(N, 1, seed)
    V = CC.ints_so(prec)
    h, f = CC.form_fock(C, h_ao, V, 2)
    B = CC.Blocks(V, f, 2)
    assert np.abs(B.fov).max() < 1e-10       # RHF reference: the Fae quirk term vanishes
    e_cc, T1, T2, it = CC.ccsd_converge(B, tol=1e-13)
    e_fci = fci.fci_two_electron(h, V)
    e_hf = h[0, 0] + h[1, 1] + V[0, 1, 0, 1]
    assert np.isclose(e_cc, e_fci - e_hf, rtol=0, atol=1e-10), (e_cc, e_fci - e_hf)
    # converged amplitudes are a fixed point of the update
    T1n, T2n = CC.update(B, T1, T2)
    assert np.abs(T1n - T1).max() < 1e-8 and np.abs(T2n - T2).max() < 1e-8


def test_ccsd_ur_driver_shape_and_convergence():
    N, nocc_sp = 5, 2
    h_ao, eri, C, F, prec = _system(N, nocc_sp, seed=11)
    r = CC.ccsd_ur(prec, C, h_ao, 2 * nocc_sp, ion_rep=1.5, maxiter=20)
    assert len(r["energies"]) == 20
    assert np.isclose(r["energies"][0], r["emp2"], rtol=1e-13)
    B = CC.Blocks(r["V"], r["f"], 2 * nocc_sp)
    e_conv, *_ = CC.ccsd_converge(B)
    assert abs(r["energies"][-1] - e_conv) < 1e-7
    # HF check reproduces the SCF energy of the closed-shell determinant
    e_scf = 0.5 * np.sum(2 * C[:nocc_sp].T @ C[:nocc_sp] * (h_ao + F)) + 1.5
    assert np.isclose(r["e_hf_check"], e_scf, rtol=1e-12)


def test_wabef_tau_term_equals_a_doubled_wmnij_tau_term():
    """The T2 contribution of Wabef's 1/4 tau_mnab <mn||ef> term through the pp
    ladder (1/2 tau_ijef W_abef) equals that of Wmnij's 1/4 tau_ijef <mn||ef>
    term through the hh ladder (1/2 tau_mnab W_mnij), for arbitrary amplitudes;
    so Wmnij with 1/2 and Wabef without its tau term give the same T2."""
    N, nocc_sp = 6, 2
    h_ao, eri, C, F, prec = _system(N, nocc_sp, seed=5)
    V = CC.ints_so(prec)
    h, f = CC.form_fock(C, h_ao, V, 2 * nocc_sp)
    B = CC.Blocks(V, f, 2 * nocc_sp)
    rng = np.random.default_rng(9)
    T1 = 0.05 * rng.standard_normal((B.o, B.v))
    T2 = 0.05 * rng.standard_normal((B.o, B.o, B.v, B.v))
    T2 = T2 - T2.transpose(1, 0, 2, 3)
    T2 = T2 - T2.transpose(0, 1, 3, 2)
    _, tau = CC.taus(T1, T2)
    wabef_tau = 0.25 * np.einsum("mnab,mnef->abef", tau, B.oovv)
    pp = 0.5 * np.einsum("ijef,abef->ijab", tau, wabef_tau)
    wmnij_tau = 0.25 * np.einsum("ijef,mnef->mnij", tau, B.oovv)
    hh = 0.5 * np.einsum("mnab,mnij->ijab", tau, wmnij_tau)
    assert np.abs(pp).max() > 1e-6
    assert np.allclose(pp, hh, rtol=0, atol=1e-15)
