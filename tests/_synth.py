"""Seeded synthetic inputs with the symmetries of real integrals (tests only)."""
import numpy as np


def eri_like(N, naux=None, seed=0):
    """8-fold symmetric, positive semidefinite (ab|cd) = sum_P X_Pab X_Pcd."""
    rng = np.random.default_rng(seed)
    naux = naux or 3 * N
    X = rng.standard_normal((naux, N, N)) / N
    X = 0.5 * (X + X.transpose(0, 2, 1))
    return np.einsum("Pab,Pcd->abcd", X, X)


def orthonormal(N, seed=0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((N, N)))
    return q


def orbital_energies(nocc, nvir, seed=0):
    """Occupied in [-1.5, -0.3], virtual in [0.1, 3.0] (hartree), sorted."""
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-1.5, -0.3, nocc))
    ev = np.sort(rng.uniform(0.1, 3.0, nvir))
    return eo, ev


def bia(naux, nocc, nvir, seed=0, scale=0.1):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal((naux, nocc, nvir))


def metric(naux, seed=0, cond=1e4):
    """SPD naux x naux metric with a prescribed condition number."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    lam = np.geomspace(1.0, cond, naux)
    return (q * lam) @ q.T


def rhf_orthonormal(h, eri, nocc_spatial, maxiter=200, tol=1e-13):
    """Closed-shell RHF in an orthonormal basis (S = I): returns (C, eps, F)
    with C[p, mu] (row = MO) as in the reference layout."""
    n = h.shape[0]
    e, Cc = np.linalg.eigh(h)
    for _ in range(maxiter):
        Co = Cc[:, :nocc_spatial]
        P = 2.0 * Co @ Co.T
        F = h + np.einsum("mnls,ls->mn", eri, P) - 0.5 * np.einsum("mlns,ls->mn", eri, P)
        e_new, C_new = np.linalg.eigh(F)
        dP = np.abs(2.0 * C_new[:, :nocc_spatial] @ C_new[:, :nocc_spatial].T - P).max()
        Cc, e = C_new, e_new
        if dP < tol:
            break
    Co = Cc[:, :nocc_spatial]
    P = 2.0 * Co @ Co.T
    F = h + np.einsum("mnls,ls->mn", eri, P) - 0.5 * np.einsum("mlns,ls->mn", eri, P)
    return Cc.T.copy(), e, F


def model_hamiltonian(N, seed=0, eri_scale=0.05):
    """Symmetric one-electron h with a gap and a small 8-fold symmetric ERI."""
    rng = np.random.default_rng(seed)
    h = np.diag(np.linspace(-2.0, 1.5, N)) + 0.05 * rng.standard_normal((N, N))
    h = 0.5 * (h + h.T)
    eri = eri_scale * eri_like(N, seed=seed + 1) * N
    return h, eri


def nonorthogonal_system(N, seed=0, spread=0.15, eri_scale=0.05, nocc=None):
    """(H, S, eri, h, eri_orth, X): the model Hamiltonian of an orthonormal
    basis phi re-expressed in the non-orthogonal basis chi_mu = sum_k phi_k X[k, mu]:
    S = X^T X, H = X^T h X, (mu nu|la si) = X^4 applied to (kl|mn).
    With nocc, h has a gap of 1.5 Eh after its nocc lowest levels, so that the
    damped SCF converges in a dozen iterations."""
    if nocc is None:
        h, eri = model_hamiltonian(N, seed=seed, eri_scale=eri_scale)
    else:
        rng = np.random.default_rng(seed)
        d = np.concatenate([np.linspace(-2.0, -1.0, nocc), np.linspace(0.5, 2.5, N - nocc)])
        h = np.diag(d) + 0.05 * rng.standard_normal((N, N))
        h = 0.5 * (h + h.T)
        eri = eri_scale * eri_like(N, seed=seed + 1) * N
    rng = np.random.default_rng(seed + 7)
    X = np.eye(N) + spread * rng.standard_normal((N, N))
    S = X.T @ X
    H = X.T @ h @ X
    eri_ao = np.einsum("km,ln,ar,bs,klab->mnrs", X, X, X, X, eri, optimize=True)
    return H, S, eri_ao, h, eri, X


def factored_system(N, nocc, seed=0, spread=0.15, eri_scale=0.05):
    """(H, S, B): the gapped model of nonorthogonal_system with its integrals
    as exact factors, (mu nu|la si) = sum_Q B[Q, mu, nu] B[Q, la, si]."""
    H, S, eri, h, eri_o, X = nonorthogonal_system(N, seed=seed, spread=spread, eri_scale=eri_scale, nocc=nocc)
    rng = np.random.default_rng(seed + 1)
    naux = 3 * N
    Xq = rng.standard_normal((naux, N, N)) / N
    Xq = 0.5 * (Xq + Xq.transpose(0, 2, 1))
    B = np.sqrt(eri_scale * N) * np.einsum("km,Qkl,ln->Qmn", X, Xq, X)
    return H, S, B
