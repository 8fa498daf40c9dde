"""Pins the RHF-side oracle (oracle/scf.py): two formulations of the Fock
build, the MO-basis HF energy, basis invariance of the converged SCF, the
converged commutator and closed-form nuclear repulsion.  No GPU needed."""
import numpy as np
import pytest

from oracle import posthf as O
from oracle import scf as SCF
from tests import _synth as S


def _density(N, nocc, seed):
    C = S.orthonormal(N, seed=seed)
    return SCF.build_pmat(C, 2 * nocc, None, 0.0), C


@pytest.mark.parametrize("N", [1, 2, 3, 5, 6])
def test_fock_list_equals_dense(N):
    """The per-class image loop of build_Fmat against the dense J/K contraction."""
    dense = S.eri_like(N, seed=10 + N)
    val, idx, sc = O.sorted_unique_list(dense)
    rng = np.random.default_rng(N)
    H = rng.standard_normal((N, N))
    H = H + H.T
    P, _ = _density(N, max(1, N // 2), seed=N)
    F_list = SCF.build_fmat_list(N, P, H, val, idx, sc)
    F_dense = SCF.fock_dense(dense, P, H)
    assert np.allclose(F_list, F_dense, rtol=0, atol=1e-13)


def test_fock_list_drop_class_detected():
    """Every class contributes: dropping one class's images changes F."""
    N = 5
    dense = S.eri_like(N, seed=3)
    val, idx, sc = O.sorted_unique_list(dense)
    P, _ = _density(N, 2, seed=4)
    H = np.zeros((N, N))
    F = SCF.build_fmat_list(N, P, H, val, idx, sc)
    for k in range(5):
        lo = 0 if k == 0 else sc[k - 1]
        hi = sc[k] if k < 4 else len(val)
        v2 = np.array(val, copy=True)
        v2[lo:hi] = 0.0
        assert not np.allclose(SCF.build_fmat_list(N, P, H, v2, idx, sc), F, atol=1e-8), k


def test_hf_energy_mo_formula():
    """1/2 tr P(H+F) equals sum_i 2h_ii + sum_ij [2(ii|jj) - (ij|ij)] in the MO basis."""
    N, nocc = 6, 2
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=5)
    T = SCF.calc_s12(Sm)
    C = S.orthonormal(N, seed=6) @ T           # rows orthonormal in the S metric
    P = SCF.build_pmat(C, 2 * nocc, None, 0.0)
    F = SCF.fock_dense(eri, P, H)
    e_ao = SCF.calc_e_el(F, P, H)
    h = C @ H @ C.T
    g = O.motrans(C, eri)
    e_mo = sum(2 * h[i, i] for i in range(nocc))
    e_mo += sum(2 * g[i, i, j, j] - g[i, j, i, j] for i in range(nocc) for j in range(nocc))
    assert abs(e_ao - e_mo) < 1e-12


def test_s12_and_diag():
    N = 7
    _, Sm, *_ = S.nonorthogonal_system(N, seed=8)
    T = SCF.calc_s12(Sm)
    assert np.allclose(T @ Sm @ T, np.eye(N), atol=1e-12)
    rng = np.random.default_rng(9)
    F = rng.standard_normal((N, N))
    F = F + F.T
    C, eps = SCF.diag_fmat(F, T)
    assert np.allclose(C @ Sm @ C.T, np.eye(N), atol=1e-11)
    assert np.allclose(C @ F @ C.T, np.diag(eps), atol=1e-11)
    assert np.all(np.diff(eps) >= 0)


@pytest.mark.parametrize("N,nocc", [(6, 2), (9, 3)])
def test_scf_basis_invariance(N, nocc):
    """Damped RHF + DIIS in a non-orthogonal basis converges to the energy of
    plain Roothaan iterations in the orthonormal basis."""
    H, Sm, eri, h, eri_o, X = S.nonorthogonal_system(N, seed=20 + N)
    r = SCF.run_scf(H, Sm, lambda P: SCF.fock_dense(eri, P, H), np.zeros((N, N)), 2 * nocc,
                    ion_rep=0.0, e_tol=1e-12)
    assert r["converged"] and r["diis"] and r["iterations"] > 8
    assert abs(sum(r["last_coeffs"]) - 1.0) < 1e-10
    Co, eps_o, Fo = S.rhf_orthonormal(h, eri_o, nocc)
    Po = 2 * Co[:nocc].T @ Co[:nocc]
    e_orth = 0.5 * np.sum(Po * (h + Fo))
    assert abs(r["total"] - e_orth) < 1e-10
    assert abs(r["energies"][-1] - e_orth) < 1e-10
    assert np.allclose(np.sort(r["eps"]), np.sort(eps_o), atol=1e-7)
    comm = SCF.calc_diis_error(r["F"], r["P"], Sm)
    assert np.abs(comm).max() < 1e-6
    # the density is idempotent in the S metric: P S P = 2 P
    assert np.allclose(r["P"] @ Sm @ r["P"], 2 * r["P"], atol=1e-8)


def test_scf_reference_threshold_and_ion_rep_offset():
    N, nocc = 6, 2
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=31)
    fock = lambda P: SCF.fock_dense(eri, P, H)   # noqa: E731
    r0 = SCF.run_scf(H, Sm, fock, np.zeros((N, N)), 2 * nocc, ion_rep=0.0)
    r1 = SCF.run_scf(H, Sm, fock, np.zeros((N, N)), 2 * nocc, ion_rep=3.25)
    assert r0["iterations"] == r1["iterations"]
    assert abs(r1["total"] - r0["total"] - 3.25) < 1e-12
    assert abs(np.diff(r0["energies"])[-1]) <= 1e-8
    r2 = SCF.run_scf(H, Sm, fock, np.zeros((N, N)), 2 * nocc, maxiter=3)
    assert r2["iterations"] == 3 and not r2["converged"]


def test_ion_rep_closed_forms():
    R = 1.4
    assert abs(SCF.calc_ion_rep(np.array([[0, 0, 0], [0, 0, R]]), np.array([1.0, 1.0])) - 1 / R) < 1e-15
    tri = np.array([[0, 0, 0], [R, 0, 0], [R / 2, R * np.sqrt(3) / 2, 0]])
    assert abs(SCF.calc_ion_rep(tri, np.array([8.0, 8.0, 8.0])) - 3 * 64 / R) < 1e-12
