"""HIP Fock build, density and RHF driver (libcorr_gpu.so) against the CPU
oracle (oracle/scf.py) and, at full sizes, a low-rank closed form.

Bars: FP64 Fock matrices to 1e-12 absolute (oracle sizes) / 1e-11 relative
(closed form), SCF energies to 1e-10 Eh per iteration, identical iteration
counts and DIIS start."""
import sys

import numpy as np
import pytest
import torch

from oracle import posthf as O
from oracle import scf as SCF
from tests import _synth as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def corr():
    import corr_amd
    return corr_amd


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to("cuda")


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _sym(rng, N):
    A = rng.standard_normal((N, N))
    return A + A.T


@pytest.mark.parametrize("N", [1, 5, 13, 64, 70])
def test_fock_dense_vs_oracle_loop(corr, N):
    say(f"fock dense N={N}")
    rng = np.random.default_rng(N)
    dense = S.eri_like(N, seed=N)
    val, idx, sc = O.sorted_unique_list(dense)
    P, H = _sym(rng, N), _sym(rng, N)
    ref = SCF.build_fmat_list(N, P, H, val, idx, sc) if N <= 13 else SCF.fock_dense(dense, P, H)
    F = corr.build_Fmat(dev(dense), dev(P), dev(H)).cpu().numpy()
    assert np.abs(F - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("N", [2, 6, 11])
def test_fock_list_vs_oracle_loop(corr, N):
    say(f"fock list N={N}")
    rng = np.random.default_rng(40 + N)
    dense = S.eri_like(N, seed=40 + N)
    val, idx, sc = O.sorted_unique_list(dense)
    P, H = _sym(rng, N), _sym(rng, N)
    ref = SCF.build_fmat_list(N, P, H, val, idx, sc)
    it = torch.from_numpy(idx.astype(np.int16)).to("cuda")
    F = corr.build_Fmat_list(N, dev(val), it, sc, dev(P), dev(H)).cpu().numpy()
    assert np.abs(F - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("N", [192, 257, 330])
def test_fock_full_size_low_rank_closed_form(corr, N):
    """(pq|rs) = sum_Q X_Q[pq] X_Q[rs] gives J = sum_Q X_Q tr(X_Q P) and
    K = sum_Q X_Q P X_Q: exact for any N, evaluated with FP64 torch GEMMs."""
    say(f"fock closed form N={N}")
    g = torch.Generator(device="cuda").manual_seed(N)
    nq = 6
    X = torch.randn(nq, N, N, device="cuda", dtype=torch.float64, generator=g) / N
    X = 0.5 * (X + X.transpose(1, 2)).contiguous()
    Xf = X.reshape(nq, N * N)
    D = (Xf.t().contiguous() @ Xf).reshape(N, N, N, N)
    P = torch.randn(N, N, device="cuda", dtype=torch.float64, generator=g)
    P = (P + P.t()).contiguous()
    H = torch.zeros(N, N, device="cuda", dtype=torch.float64)
    F = corr.build_Fmat(D, P, H)
    F2 = corr.build_Fmat(D, P, H)
    assert torch.equal(F, F2)                                   # deterministic
    tr = (Xf @ P.reshape(-1))                                   # tr(X_Q P)
    Jr = (tr[:, None] * Xf).sum(0).reshape(N, N)
    Kr = torch.einsum("qab,bc,qcd->ad", X, P, X)
    ref = Jr - 0.5 * Kr
    err = (F - ref).abs().max().item()
    assert err < 1e-11 * ref.abs().max().item(), err
    del D, F, F2
    torch.cuda.empty_cache()


def test_build_pmat_damped(corr):
    say("build_pmat")
    N, nroe = 9, 6
    C = S.orthonormal(N, seed=5)
    rng = np.random.default_rng(6)
    Pold = _sym(rng, N)
    P0 = corr.build_Pmat(dev(C), nroe).cpu().numpy()
    assert np.abs(P0 - SCF.build_pmat(C, nroe, None, 0.0)).max() < 1e-14
    Pd = dev(Pold)
    corr.build_Pmat(dev(C), nroe, 0.5, Pd)
    assert np.abs(Pd.cpu().numpy() - SCF.build_pmat(C, nroe, Pold, 0.5)).max() < 1e-14


@pytest.mark.parametrize("N,nocc,guess", [(9, 3, "core"), (24, 5, "core"), (24, 5, "given")])
def test_run_scf_vs_oracle(corr, N, nocc, guess):
    say(f"run_scf N={N} nocc={nocc} guess={guess}")
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=60 + N, nocc=nocc)
    fock = lambda P: SCF.fock_dense(eri, P, H)   # noqa: E731
    if guess == "core":
        C0 = np.zeros((N, N))
    else:
        C0 = SCF.diag_fmat(H + 0.1 * np.eye(N), SCF.calc_s12(Sm))[0]
    ref = SCF.run_scf(H, Sm, fock, SCF.check_initial_guess(C0, H, SCF.calc_s12(Sm))[0], 2 * nocc, ion_rep=1.75)
    r = corr.run_scf(dev(H), dev(Sm), dev(eri), dev(C0), 2 * nocc, ion_rep=1.75)
    assert r["core_guess"] == (1 if guess == "core" else 0)
    assert r["iterations"] == ref["iterations"] and r["converged"] == 1
    assert bool(r["diis"]) == ref["diis"]
    assert np.abs(r["energies"] - ref["energies"]).max() < 1e-10
    assert abs(r["e_total"] - ref["total"]) < 1e-10
    assert abs(r["e_one"] - ref["one"]) < 1e-10 and abs(r["e_two"] - ref["two"]) < 1e-10
    assert np.abs(r["MOens"].cpu().numpy() - ref["eps"]).max() < 1e-9
    assert np.abs(r["Pmat"].cpu().numpy() - ref["P"]).max() < 1e-9
    assert np.abs(r["Fmat"].cpu().numpy() - ref["F"]).max() < 1e-9


def test_run_scf_tight_matches_orthonormal_rhf(corr):
    """Basis invariance on the GPU path: e_tol 1e-12 in the non-orthogonal basis
    against plain Roothaan iterations in the orthonormal basis."""
    say("run_scf tight")
    N, nocc = 20, 4
    H, Sm, eri, h, eri_o, _ = S.nonorthogonal_system(N, seed=77, nocc=nocc)
    r = corr.run_scf(dev(H), dev(Sm), dev(eri), dev(np.zeros((N, N))), 2 * nocc, e_tol=1e-12)
    Co, eps_o, Fo = S.rhf_orthonormal(h, eri_o, nocc)
    Po = 2 * Co[:nocc].T @ Co[:nocc]
    assert abs(r["e_total"] - 0.5 * np.sum(Po * (h + Fo))) < 1e-10
    assert r["diis"] == 1 and r["diis_singular"] == 0


def test_run_scf_rejects_indefinite_overlap(corr):
    say("run_scf indefinite S")
    N = 6
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=3)
    Sbad = Sm.copy()
    Sbad[0, 0] = -1.0
    with pytest.raises(corr.CorrError) as e:
        corr.run_scf(dev(H), dev(Sbad), dev(eri), dev(np.zeros((N, N))), 4)
    assert e.value.status == 5
    r = corr.run_scf(dev(H), dev(Sm), dev(eri), dev(np.zeros((N, N))), 4)   # context still usable
    assert r["converged"] == 1


def test_run_scf_maxiter_cap(corr):
    say("run_scf maxiter 3")
    N, nocc = 9, 3
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=5, nocc=nocc)
    C0 = SCF.check_initial_guess(np.zeros((N, N)), H, SCF.calc_s12(Sm))[0]
    ref = SCF.run_scf(H, Sm, lambda P: SCF.fock_dense(eri, P, H), C0, 2 * nocc, maxiter=3)
    r = corr.run_scf(dev(H), dev(Sm), dev(eri), dev(np.zeros((N, N))), 2 * nocc, maxiter=3)
    assert r["iterations"] == 3 and r["converged"] == 0 and not ref["converged"]
    assert np.abs(r["energies"] - ref["energies"]).max() < 1e-10
    assert abs(r["e_total"] - ref["total"]) < 1e-10


def test_fock_ri_equals_dense_for_factorised_integrals(corr):
    """With (mu nu|la si) = sum_Q B^Q_mu nu B^Q_la si exactly, the RI-JK Fock
    matrix equals the dense Fock build, and the two SCF drivers agree."""
    say("fock RI-JK vs dense")
    N, nocc = 40, 6
    rng = np.random.default_rng(11)
    H, Sm, B = S.factored_system(N, nocc, seed=12)
    eri = np.einsum("Qab,Qcd->abcd", B, B)
    P, Hr = _sym(rng, N), _sym(rng, N)
    Fd = corr.build_Fmat(dev(eri), dev(P), dev(Hr)).cpu().numpy()
    Fr = corr.build_Fmat_ri(dev(B), dev(P), dev(Hr)).cpu().numpy()
    assert np.abs(Fd - Fr).max() < 1e-12 * np.abs(Fd).max()
    assert np.abs(Fr - SCF.fock_dense(eri, P, Hr)).max() < 1e-12 * np.abs(Fd).max()
    C0 = np.zeros((N, N))
    rd = corr.run_scf(dev(H), dev(Sm), dev(eri), dev(C0), 2 * nocc, ion_rep=0.5)
    rr = corr.run_scf_ri(dev(H), dev(Sm), dev(B), dev(C0), 2 * nocc, ion_rep=0.5)
    assert rd["converged"] == 1 and rr["iterations"] == rd["iterations"]
    assert np.abs(rr["energies"] - rd["energies"]).max() < 1e-10
    assert abs(rr["e_total"] - rd["e_total"]) < 1e-10
