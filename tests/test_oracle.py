"""Pins the CPU oracle (oracle/posthf.py) against closed forms, brute force,
invariants and finite differences.  No GPU needed."""
import numpy as np
import pytest

from oracle import posthf as O
from tests import _synth as S


def test_build_fmo_identity_and_trace():
    rng = np.random.default_rng(1)
    F = rng.standard_normal((7, 7))
    F = F + F.T
    assert np.array_equal(O.build_fmo(np.eye(7), F), F)
    C = S.orthonormal(7, seed=2)
    assert np.isclose(np.trace(O.build_fmo(C, F)), np.trace(F), atol=1e-12)
    # explicit quadruple loop of motrans.cc:23-31 on a small case
    C = rng.standard_normal((4, 4))
    F = rng.standard_normal((4, 4))
    ref = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            for m in range(4):
                for n in range(4):
                    ref[i, j] += C[i, m] * F[m, n] * C[j, n]
    assert np.allclose(O.build_fmo(C, F), ref, atol=1e-13)


@pytest.mark.parametrize("N", [1, 2, 5, 7])
def test_unique_list_roundtrip(N):
    dense = S.eri_like(N, seed=N)
    val, idx, sc = O.sorted_unique_list(dense)
    M = N * (N + 1) // 2
    assert len(val) == M * (M + 1) // 2           # no screening on dense random data
    assert sc[0] == N                              # (aa|aa)
    assert sc[1] - sc[0] == N * (N - 1) // 2       # (aa|cc), a > c
    assert sc[2] - sc[1] == N * (N - 1) // 2       # (ab|ab), a > b
    back = O.unfold_sorted(N, val, idx, sc)
    assert np.array_equal(back, dense)


def test_unique_list_images_never_collide():
    # every dense element is produced exactly once
    N = 6
    dense = S.eri_like(N, seed=3)
    val, idx, sc = O.sorted_unique_list(dense)
    counts = O.unfold_sorted(N, np.ones_like(val), idx, sc)
    assert np.array_equal(counts, np.ones((N,) * 4))


def test_unique_list_screening_drops_small():
    N = 4
    dense = S.eri_like(N, seed=4)
    dense[np.abs(dense) < 2e-3] = 0.0
    val, idx, sc = O.sorted_unique_list(dense)
    assert np.all(np.abs(val) > 1e-12)
    assert np.array_equal(O.unfold_sorted(N, val, idx, sc), dense)


def test_motrans_bruteforce_and_identity():
    N = 5
    dense = S.eri_like(N, seed=5)
    val, idx, sc = O.sorted_unique_list(dense)
    rng = np.random.default_rng(6)
    C = rng.standard_normal((N, N))
    mo = O.motrans(C, dense)
    for (i, j, k, l) in [(0, 0, 0, 0), (1, 2, 3, 4), (4, 3, 2, 1), (2, 2, 0, 4), (3, 1, 3, 1)]:
        assert np.isclose(mo[i, j, k, l], O.calc_mo2int_bf(i, j, k, l, C, val, idx, sc),
                          rtol=1e-12, atol=1e-13)
    assert np.allclose(O.motrans(np.eye(N), dense), dense, atol=1e-15)
    # 8-fold symmetry of the transformed integrals
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.allclose(mo, mo.transpose(perm), atol=1e-12)
    # a different contraction order gives the same tensor
    alt = np.einsum("ia,jb,kc,ld,abcd->ijkl", C, C, C, C, dense, optimize=True)
    assert np.allclose(mo, alt, atol=1e-12)


def test_canonical_mp2_two_level_closed_form():
    # one occupied, one virtual: EMP2_OS = -(ia|ia)^2 / (2(e_a - e_i)), EMP2_SS = 0
    N, o = 2, 1
    I4 = np.zeros((2, 2, 2, 2))
    I4[0, 1, 0, 1] = 0.3
    F = np.diag([-0.5, 0.7])
    e = O.canonical_mp2(I4, F, o)
    assert np.isclose(e["EMP2_OS"], -0.09 / 2.4, rtol=1e-15)
    assert e["EMP2_SS"] == 0.0
    assert np.isclose(e["EMP2_SCS"], 1.2 * e["EMP2_OS"])


def test_ri_equals_canonical_when_factorised():
    naux, o, v = 11, 3, 5
    B = S.bia(naux, o, v, seed=7)
    eo, ev = S.orbital_energies(o, v, seed=8)
    N = o + v
    I4 = np.zeros((N, N, N, N))
    I4[:o, o:, :o, o:] = np.einsum("Qia,Qjb->iajb", B, B)
    FMo = np.diag(np.concatenate([eo, ev]))
    can = O.canonical_mp2(I4, FMo, o)
    ri = O.rimp2(B, eo, ev)
    for k in can:
        assert np.isclose(ri[k], can[k], rtol=1e-13, atol=1e-15)
    assert np.isclose(O.rimp2_dense(B, eo, ev)["EMP2"], ri["EMP2"], rtol=1e-13)


def test_ri_metric_eigen_cholesky_cutoff():
    naux = 30
    J = S.metric(naux, seed=9, cond=1e6)
    M, kept = O.ri_metric_eigen(J)
    assert kept == naux
    assert np.allclose(M @ J @ M, np.eye(naux), atol=1e-8)
    L = O.ri_metric_cholesky(J)
    assert np.allclose(L.T @ L, np.linalg.inv(J), rtol=1e-8, atol=1e-8)
    X = np.random.default_rng(10).standard_normal((naux, 17))
    assert np.allclose((M @ X).T @ (M @ X), (L @ X).T @ (L @ X), rtol=1e-9, atol=1e-9)
    # rank-deficient: eigenvalues below 1e-10 * l_max are dropped
    q, _ = np.linalg.qr(np.random.default_rng(11).standard_normal((naux, naux)))
    lam = np.geomspace(1.0, 10.0, naux)
    lam[:4] = [1e-14, -1e-13, 5e-12, 2e-11]   # 4 below the 1e-10 relative cutoff
    Jd = (q * lam) @ q.T
    Md, kept = O.ri_metric_eigen(Jd)
    assert kept == naux - 4
    P = Md @ Jd @ Md   # projector onto the kept space
    assert np.allclose(P @ P, P, atol=1e-8)
    assert np.isclose(np.trace(P), naux - 4, atol=1e-8)


def test_transform_ri_explicit_loops():
    naux, N, o = 3, 4, 2
    rng = np.random.default_rng(12)
    C = rng.standard_normal((N, N))
    BPQ = rng.standard_normal((naux, N, N))
    ref = np.zeros((naux, o, N - o))
    for Q in range(naux):
        for i in range(o):
            for a in range(N - o):
                for m in range(N):
                    for n in range(N):
                        ref[Q, i, a] += C[i, m] * C[o + a, n] * BPQ[Q, m, n]
    assert np.allclose(O.transform_ri(C, BPQ, o), ref, atol=1e-13)


@pytest.mark.parametrize("c_os,c_ss", [(1.0, 1.0), (1.2, 1.0 / 3.0)])
def test_rimp2_gradient_finite_differences(c_os, c_ss):
    naux, o, v = 7, 3, 4
    B = S.bia(naux, o, v, seed=13)
    eo, ev = S.orbital_energies(o, v, seed=14)

    def energy(B_, eo_, ev_):
        e = O.rimp2(B_, eo_, ev_)
        return c_os * e["EMP2_OS"] + c_ss * e["EMP2_SS"]

    E, dB, deo, dev = O.rimp2_gradient(B, eo, ev, c_os, c_ss)
    assert np.isclose(E, energy(B, eo, ev), rtol=1e-13)
    h = 1e-5
    rng = np.random.default_rng(15)
    for _ in range(12):
        q, i, a = rng.integers(naux), rng.integers(o), rng.integers(v)
        Bp, Bm = B.copy(), B.copy()
        Bp[q, i, a] += h
        Bm[q, i, a] -= h
        fd = (energy(Bp, eo, ev) - energy(Bm, eo, ev)) / (2 * h)
        assert np.isclose(dB[q, i, a], fd, rtol=1e-6, atol=1e-10)
    for i in range(o):
        ep, em = eo.copy(), eo.copy()
        ep[i] += h
        em[i] -= h
        assert np.isclose(deo[i], (energy(B, ep, ev) - energy(B, em, ev)) / (2 * h), rtol=1e-6, atol=1e-10)
    for a in range(v):
        ep, em = ev.copy(), ev.copy()
        ep[a] += h
        em[a] -= h
        assert np.isclose(dev[a], (energy(B, eo, ep) - energy(B, eo, em)) / (2 * h), rtol=1e-6, atol=1e-10)


def test_non_canonical_mp2_reference_loops_small():
    """Vectorised restatement against the reference's explicit loops
    (loc_mp2.cc:24-90) on a tiny non-canonical system, 3 iterations."""
    N, o = 5, 2
    v = N - o
    rng = np.random.default_rng(3)
    prec = O.motrans(S.orthonormal(N, seed=4), S.eri_like(N, seed=5))
    F = np.diag(np.concatenate([[-1.2, -0.7], [0.3, 0.8, 1.4]]))
    F[0, 1] = F[1, 0] = 0.05
    E, Rs, _ = O.non_canonical_mp2(prec, F, o, iters=3)
    K = lambda i, j, a, b: prec[i, a + o, j, b + o]   # noqa: E731
    Dn = lambda i, j, a, b: F[o + a, o + a] + F[o + b, o + b] - F[i, i] - F[j, j]   # noqa: E731
    T = np.zeros((o, o, v, v))
    for i, j, a, b in np.ndindex(o, o, v, v):
        T[i, j, a, b] = -K(i, j, a, b) / Dn(i, j, a, b)
    for it in range(3):
        G = np.zeros_like(T)
        for i, j, a, b in np.ndindex(o, o, v, v):
            for k in range(o):
                if k != i:
                    G[i, j, a, b] += -F[i, k] * T[k, j, a, b]
        R = np.zeros_like(T)
        rsum = 0.0
        for i, j, a, b in np.ndindex(o, o, v, v):
            R[i, j, a, b] = K(i, j, a, b) + Dn(i, j, a, b) * T[i, j, a, b] + G[i, j, a, b] + G[j, i, b, a]
            rsum += abs(R[i, j, a, b])
        e = sum((K(i, j, a, b) + R[i, j, a, b]) * (2 * T[i, j, a, b] - T[i, j, b, a])
                for i, j, a, b in np.ndindex(o, o, v, v))
        for i, j, a, b in np.ndindex(o, o, v, v):
            T[i, j, a, b] += -R[i, j, a, b] / Dn(i, j, a, b)
        assert abs(E[it] - e) < 1e-14 and abs(Rs[it] - rsum) < 1e-13


def test_non_canonical_mp2_canonical_limit_and_occupied_invariance():
    """Diagonal Fock: iteration 0 is canonical MP2 with Rsum = 0.  Rotating
    the occupied orbitals makes the occupied Fock block non-diagonal; the
    iterations converge back to the canonical energy."""
    N, o = 10, 3
    h, eri = S.model_hamiltonian(N, seed=21)
    C, eps, F = S.rhf_orthonormal(h, eri, o)
    prec = O.motrans(C, eri)
    FMo = O.build_fmo(C, F)
    e_can = O.canonical_mp2(prec, FMo, o)["EMP2"]
    E, Rs, _ = O.non_canonical_mp2(prec, FMo, o, iters=2)
    assert abs(E[0] - e_can) < 1e-13 and Rs[0] < 1e-12
    U = S.orthonormal(o, seed=22)
    C2 = C.copy()
    C2[:o] = U @ C[:o]
    prec2 = O.motrans(C2, eri)
    FMo2 = O.build_fmo(C2, F)
    assert np.abs(FMo2[:o, :o] - np.diag(np.diag(FMo2[:o, :o]))).max() > 1e-3
    E2, Rs2, _ = O.non_canonical_mp2(prec2, FMo2, o, iters=20)
    assert abs(E2[-1] - e_can) < 1e-10 and Rs2[-1] < Rs2[0] * 1e-6
