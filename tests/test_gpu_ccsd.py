"""GPU parity of the spin-orbital CCSD against the CPU oracle (oracle/ccsd.py).

FP64: amplitudes to 1e-12 absolute, energies to 1e-10 Eh.  FP32 MFMA
contractions: energies to 1e-6 Eh on these systems."""
import sys

import numpy as np
import pytest
import torch

from oracle import ccsd as CC
from oracle import fci
from oracle import posthf as O
from tests import _synth as S

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def corr():
    import corr_amd
    return corr_amd


def system(N, nocc_sp, seed, canonical=True):
    h_ao, eri = S.model_hamiltonian(N, seed=seed)
    C, eps, F = S.rhf_orthonormal(h_ao, eri, nocc_sp)
    if not canonical:
        C = S.orthonormal(N, seed=seed + 7)
    return h_ao, C, O.motrans(C, eri)


def test_ints_so(corr):
    _say("ccsd ints_so")
    prec = S.eri_like(5, seed=1)
    got = corr.ints_so(dev(prec)).cpu().numpy()
    assert np.array_equal(got, CC.ints_so(prec))


@pytest.mark.parametrize("N,nocc_sp,canonical", [(4, 1, True), (6, 2, True), (6, 2, False), (9, 3, True)])
def test_update_and_energy(corr, N, nocc_sp, canonical):
    _say(f"ccsd update N={N} canonical={canonical}")
    h_ao, C, prec = system(N, nocc_sp, seed=N, canonical=canonical)
    V = CC.ints_so(prec)
    h, f = CC.form_fock(C, h_ao, V, 2 * nocc_sp)
    B = CC.Blocks(V, f, 2 * nocc_sp)
    T1, T2, _ = CC.guess(B)
    rng = np.random.default_rng(N)
    T1 = T1 + 0.01 * rng.standard_normal(T1.shape)     # non-zero T1 exercises every term
    r1, r2 = CC.update(B, T1, T2)
    g1, g2 = corr.ccsd_update(dev(V), dev(f), dev(T1), dev(T2))
    assert np.abs(g1.cpu().numpy() - r1).max() < 1e-12 * max(1.0, np.abs(r1).max())
    assert np.abs(g2.cpu().numpy() - r2).max() < 1e-12 * max(1.0, np.abs(r2).max())
    e = corr.ccsd_energy(dev(V), dev(f), dev(T1), dev(T2))
    assert abs(e - CC.energy(B, T1, T2)) < 1e-12
    f1, f2 = corr.ccsd_update(dev(V), dev(f), dev(T1), dev(T2), prec="fp32")
    assert np.abs(f2.cpu().numpy() - r2).max() < 1e-5 * max(1.0, np.abs(r2).max())


def test_ccsd_ur_reference_mode(corr):
    N, nocc_sp = 7, 2
    _say("ccsd_ur reference N=7")
    h_ao, C, prec = system(N, nocc_sp, seed=21)
    ref = CC.ccsd_ur(prec, C, h_ao, 2 * nocc_sp, ion_rep=2.5, maxiter=20)
    got = corr.ccsd_ur(dev(prec), dev(C), dev(h_ao), 2 * nocc_sp, 2.5, maxiter=20)
    assert got["iterations"] == 20
    assert np.allclose(got["energies"], ref["energies"], rtol=0, atol=1e-10)
    assert abs(got["emp2"] - ref["emp2"]) < 1e-12
    assert abs(got["e_hf_check"] - ref["e_hf_check"]) < 1e-10
    assert np.abs(got["T2"].cpu().numpy() - ref["T2"]).max() < 1e-11


def test_ccsd_diis_converges_to_oracle(corr):
    """DIIS to the oracle's root.  This system's spin-flip sector is unstable
    (plain iterations diverge; rounding noise in spin-forbidden amplitudes grows
    to 0.06 and DIIS lands on a spin-broken root), so it also checks that the
    spin-blocked ring terms hand over to the dense ones once forbidden
    amplitudes appear: without that the run stops 1e-5 Eh away, at amplitudes
    that are not a root (oracle residual 4.5e-3)."""
    N, nocc_sp = 7, 2
    _say("ccsd diis N=7")
    h_ao, C, prec = system(N, nocc_sp, seed=23)
    V = CC.ints_so(prec)
    h, f = CC.form_fock(C, h_ao, V, 2 * nocc_sp)
    e_ref, *_ = CC.ccsd_converge(CC.Blocks(V, f, 2 * nocc_sp), tol=1e-12)
    got = corr.ccsd_ur(dev(prec), dev(C), dev(h_ao), 2 * nocc_sp, 0.0, maxiter=100, mode="diis",
                       e_tol=1e-10, r_tol=1e-8)
    assert got["converged"]
    assert abs(got["e_ccsd"] - e_ref) < 1e-9
    g32 = corr.ccsd_ur(dev(prec), dev(C), dev(h_ao), 2 * nocc_sp, 0.0, maxiter=100, mode="diis",
                       e_tol=1e-10, r_tol=1e-6, prec="fp32")
    assert abs(g32["e_ccsd"] - e_ref) < 1e-6


def test_two_electron_ccsd_equals_fci(corr):
    N = 5
    _say("ccsd fci N=5")
    h_ao, C, prec = system(N, 1, seed=6)
    V = CC.ints_so(prec)
    h, f = CC.form_fock(C, h_ao, V, 2)
    e_fci = fci.fci_two_electron(h, V)
    e_hf = h[0, 0] + h[1, 1] + V[0, 1, 0, 1]
    got = corr.ccsd_ur(dev(prec), dev(C), dev(h_ao), 2, 0.0, maxiter=200, mode="diis", e_tol=1e-12,
                       r_tol=1e-10)
    assert abs(got["e_ccsd"] - (e_fci - e_hf)) < 1e-9
