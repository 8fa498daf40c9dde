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
