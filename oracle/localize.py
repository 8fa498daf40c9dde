"""Orbital localisation (test oracle: only tests/, __graft_entry__.smoke() and
bench.py's CPU leg use it).

Restates syshfwplug/pymodule.py:75-167:

  calculate_PM (pymodule.py:143-153)   Pipek-Mezey spread sum_A sum_i P_Ai^2
      with Mulliken populations P_Ai = sum_{k on A} L_ki (S L)_ki
  calculate_boys (pymodule.py:156-167) Boys spread sum_d sum_i (L^T D_d L)_ii^2
      with the AO dipole matrices D_x, D_y, D_z
  localize_PM / localize_boys (pymodule.py:75-124): maximise the spread over
      U = expm(X), X antisymmetric, by scipy BFGS on 1/spread; L = C U.
      PM starts from random X (np.random.randint(0, 20, .)/100), Boys from
      X = 0.

Quirk of the reference: it sizes X as int(len(para)/2) x int(len(para)/2)
with len(para) = n(n-1)/2, which equals n only for n = 5 (one water); for
other n the assignment X[triu] = para fails.  `localize_bfgs` sizes X as n.

Both spreads are f = sum_K sum_i (Q_K)_ii^2 over symmetric matrices
Q_K = L^T M_K L: per atom A, M_A = (S P_A + P_A S)/2 with P_A the projector on
A's functions (its diagonal is P_Ai); for Boys, M_d = D_d.  `jacobi_localize`
maximises f by 2x2 rotations (Pipek & Mezey, J. Chem. Phys. 90, 4916 (1989),
eqs. 9-13): for a pair (s, t), with A = sum_K [Q_st^2 - (Q_ss - Q_tt)^2 / 4]
and B = sum_K Q_st (Q_ss - Q_tt), the angle 4g = atan2(B, -A) raises f by
A + sqrt(A^2 + B^2).

Pins (tests/test_oracle_localize.py): Mulliken populations of every
orbital sum to 1; f is invariant under orbital order and sign; the Jacobi
angle is the maximum of a brute-force scan of the 2x2 rotation; the Jacobi
and the reference's BFGS reach the same maximum on H2O (n = 5); at the
maximum every pair gradient B vanishes; Boys' trace sum_i (L^T D L)_ii is
rotation invariant.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg
import scipy.optimize


def pm_matrices(S: np.ndarray, L: np.ndarray, bf_atom: np.ndarray, natom: int) -> np.ndarray:
    """Q_A = (L^T S P_A L + L^T P_A S L) / 2 for every atom A (L: nbf x n)."""
    SL = S @ L
    Q = np.zeros((natom, L.shape[1], L.shape[1]))
    for A in range(natom):
        k = bf_atom == A
        X = L[k].T @ SL[k]
        Q[A] = 0.5 * (X + X.T)
    return Q


def calculate_pm(S: np.ndarray, L: np.ndarray, bf_atom: np.ndarray) -> float:
    """pymodule.py:143-153: PA[j][i] += L[k][i] * LS[k][i] over k on atom j."""
    LS = S @ L
    natom = int(bf_atom.max()) + 1
    PA = np.zeros((natom, L.shape[1]))
    for k in range(L.shape[0]):
        PA[bf_atom[k]] += L[k] * LS[k]
    return float(np.sum(PA ** 2))


def populations(S: np.ndarray, L: np.ndarray, bf_atom: np.ndarray) -> np.ndarray:
    LS = S @ L
    natom = int(bf_atom.max()) + 1
    PA = np.zeros((natom, L.shape[1]))
    for k in range(L.shape[0]):
        PA[bf_atom[k]] += L[k] * LS[k]
    return PA


def calculate_boys(D: np.ndarray, L: np.ndarray) -> float:
    """pymodule.py:156-167: sum over x, y, z of sum_i (L^T D_d L)_ii^2."""
    return float(sum(np.sum(np.diag(L.T @ D[d] @ L) ** 2) for d in range(3)))


def _unitary(para: np.ndarray, n: int) -> np.ndarray:
    X = np.zeros((n, n))
    X[np.triu_indices(n, 1)] = para
    return scipy.linalg.expm(X - X.T)


def localize_bfgs(C: np.ndarray, spread, seed: int | None = 0, random_start: bool = True):
    """pymodule.py:75-124: BFGS on 1/spread(C expm(X)); X sized n (see the
    module docstring for the reference's sizing quirk).  Returns (result, L)."""
    n = C.shape[1]
    npar = n * (n - 1) // 2
    if random_start:
        rng = np.random.default_rng(seed)
        U0 = rng.integers(0, 20, npar) / 100.0
    else:
        U0 = np.zeros(npar)
    res = scipy.optimize.minimize(lambda p: 1.0 / spread(C @ _unitary(p, n)), U0, method="BFGS",
                                  options={"gtol": 1e-10, "maxiter": 10000})
    return res, C @ _unitary(res.x, n)


def spread_of(Q: np.ndarray) -> float:
    return float(sum(np.sum(np.diag(q) ** 2) for q in Q))


def pair_terms(Q: np.ndarray, s: int, t: int):
    d = Q[:, s, s] - Q[:, t, t]
    q = Q[:, s, t]
    return float(np.sum(q * q - 0.25 * d * d)), float(np.sum(q * d))


def rotate(Q: np.ndarray, L: np.ndarray, s: int, t: int, g: float):
    """Orbitals s' = c s + sn t, t' = -sn s + c t (columns of L); Q <- J^T Q J."""
    c, sn = np.cos(g), np.sin(g)
    J = np.eye(L.shape[1])
    J[s, s], J[t, s], J[s, t], J[t, t] = c, sn, -sn, c
    return np.einsum("ki,Kkl,lj->Kij", J, Q, J), L @ J


def jacobi_localize(Q: np.ndarray, L: np.ndarray, tol: float = 1e-14, max_sweeps: int = 200):
    """Cyclic Jacobi sweeps over all pairs (s < t) until a sweep raises f by
    less than tol * f.  Returns (L, Q, f, sweeps)."""
    n = L.shape[1]
    f = spread_of(Q)
    for sweep in range(1, max_sweeps + 1):
        gain = 0.0
        for s in range(n):
            for t in range(s + 1, n):
                A, B = pair_terms(Q, s, t)
                g = 0.25 * np.arctan2(B, -A)
                gain += A + np.hypot(A, B)
                Q, L = rotate(Q, L, s, t, g)
        f = spread_of(Q)
        if gain <= tol * max(f, 1.0):
            return L, Q, f, sweep
    return L, Q, f, max_sweeps
