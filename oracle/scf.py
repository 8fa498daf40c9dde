"""numpy FP64 restatement of the reference's RHF input side (test oracle).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
Each function names the reference lines (Masterluke87/corr) it follows.
Pins (tests/test_oracle_scf.py; none of them retypes a formula below):
  build_fmat_list  two formulations: the per-class image loop of the reference
                   against the dense J/K contraction of the unfolded tensor;
                   the HF energy 1/2 tr P(H+F) against the MO-basis formula
                   sum_i 2h_ii + sum_ij [2(ii|jj) - (ij|ij)].
  calc_s12         S^-1/2 S S^-1/2 = I.
  diag_fmat        C S C^T = I and C F C^T = diag(eps).
  run_scf          converged total energy is basis-invariant: S != I through a
                   random non-orthogonal basis against plain Roothaan
                   iterations in the orthonormal basis; the converged
                   commutator FPS - SPF vanishes; TOTAL E = (one+two)/2 + V_nn.
  calc_ion_rep     closed forms (H2: 1/R, equilateral triangle: 3 Z^2/R).
"""
from __future__ import annotations

import numpy as np

from .posthf import entry_class


def calc_ion_rep(coord: np.ndarray, charges: np.ndarray) -> float:
    """sum_{x<y} Z_x Z_y / r_xy (calc_ion_rep, ops_rhf.cc:283-292)."""
    e = 0.0
    n = len(charges)
    for x in range(n):
        for y in range(x + 1, n):
            e += charges[x] * charges[y] / np.linalg.norm(coord[x] - coord[y])
    return float(e)


def build_pmat(C: np.ndarray, nroe: int, P_old: np.ndarray | None, damp: float) -> np.ndarray:
    """P = 2(1-damp) C_occ^T C_occ + damp P_old (build_Pmat, scf.cc:44-61:
    column-major dgemm "N","T" over the first nroe/2 rows of C, then daxpy)."""
    nocc = nroe // 2
    Co = C[:nocc]
    P = 2.0 * (1.0 - damp) * (Co.T @ Co)
    if damp != 0.0:
        P = P + damp * P_old
    return P


# (J target, J density, K target, K density) index positions per image, as the
# perm_* functions write them (ops_rhf.cc:466-547).  Image k of an entry
# (a,b,c,d) adds  F[t0,t1] += P[d0,d1] v  and  F[k0,k1] -= 0.5 P[e0,e1] v.
_A, _B, _C, _D = range(4)
_IMG = {
    1: ((_A, _B, _C, _D), (_A, _C, _B, _D)),
    2: ((_A, _B, _D, _C), (_A, _D, _B, _C)),
    3: ((_B, _A, _C, _D), (_B, _C, _A, _D)),
    4: ((_B, _A, _D, _C), (_B, _D, _A, _C)),
    5: ((_C, _D, _A, _B), (_C, _A, _D, _B)),
    6: ((_D, _C, _A, _B), (_D, _A, _C, _B)),
    7: ((_C, _D, _B, _A), (_C, _B, _D, _A)),
    8: ((_D, _C, _B, _A), (_D, _B, _C, _A)),
}
# class -> images applied (perm_1, perm_15, perm_1234, perm_1256, perm_all;
# build_Fmat, ops_rhf.cc:549-585)
_CLASS_IMAGES = {0: (1,), 1: (1, 5), 2: (1, 2, 3, 4), 3: (1, 2, 5, 6), 4: (1, 2, 3, 4, 5, 6, 7, 8)}


def build_fmat_list(N: int, P: np.ndarray, H: np.ndarray, intval, intnums, sortcount) -> np.ndarray:
    """F = H + the J and K images of every sorted unique entry
    (build_Fmat, ops_rhf.cc:549-585, with perm_* at :466-547)."""
    F = np.array(H, dtype=np.float64, copy=True)
    idx = np.asarray(intnums, dtype=np.int64).reshape(-1, 4)
    val = np.asarray(intval, dtype=np.float64)
    cls = entry_class(sortcount, len(val))
    for k, images in _CLASS_IMAGES.items():
        m = cls == k
        ix, v = idx[m], val[m]
        for im in images:
            (j0, j1, d0, d1), (k0, k1, e0, e1) = _IMG[im]
            np.add.at(F, (ix[:, j0], ix[:, j1]), P[ix[:, d0], ix[:, d1]] * v)
            np.add.at(F, (ix[:, k0], ix[:, k1]), -0.5 * P[ix[:, e0], ix[:, e1]] * v)
    return F


def fock_dense(D: np.ndarray, P: np.ndarray, H: np.ndarray) -> np.ndarray:
    """H + sum_rs (pq|rs) P_rs - 1/2 sum_qs (pq|rs) P_qs over a dense tensor."""
    N = H.shape[0]
    D2 = D.reshape(N * N, N * N)
    J = (D2 @ P.reshape(-1)).reshape(N, N)
    K = np.einsum("pqrs,qs->pr", D, P, optimize=True)
    return H + J - 0.5 * K


def calc_s12(S: np.ndarray) -> np.ndarray:
    """S^-1/2 = V diag(1/sqrt(l)) V^T, no cutoff (calc_S12, ops_rhf.cc:342-373)."""
    lam, V = np.linalg.eigh(S)
    return (V / np.sqrt(lam)) @ V.T


def diag_fmat(F: np.ndarray, Som12: np.ndarray):
    """(C, eps): F' = T F T, F' = V eps V^T, C = V^T T with rows = MOs
    (diag_Fmat, ops_rhf.cc:425-430; symmortho_mat, transform_MOs, ops_mat.cc:53-87)."""
    Fp = Som12 @ F @ Som12
    eps, V = np.linalg.eigh(Fp)
    return V.T @ Som12.T, eps


def calc_diis_error(F, P, S) -> np.ndarray:
    """S P F - F P S (calc_diis_error, scf.cc:105-141)."""
    return S @ (P @ F) - F @ (P @ S)


def calc_e_el(F, P, H) -> float:
    """1/2 sum_xy P_xy (F_yx + H_yx) (Calc_e_el, ops_rhf.cc:593-601)."""
    return 0.5 * float(np.sum(P * (F.T + H.T)))


def check_initial_guess(C, H, Som12):
    """Core guess when |trace C| < 0.1 (check_initial_guess, ops_rhf.cc:15-37)."""
    if abs(np.trace(C)) < 0.1:
        return diag_fmat(H, Som12)[0], True
    return C, False


def run_scf(H, S, fock, C0, nroe, ion_rep=0.0, maxiter=100, e_tol=1e-8, diis_size=8, diis_thresh=1e-3,
            damp0=0.5):
    """Damped RHF with DIIS(8) (run_scf, scf.cc:145-320).

    fock(P) returns the Fock matrix for a density P.  The DIIS ring starts
    zeroed and rotates every iteration; extrapolation starts (and damping
    stops) once the mean |error| < diis_thresh at iteration >= diis_size;
    the loop ends when |dE| <= e_tol or after maxiter iterations.  After the
    loop P and F are rebuilt from the last orbitals with the last damping."""
    N = H.shape[0]
    Som12 = calc_s12(S)
    C = np.array(C0, dtype=np.float64, copy=True)
    P = build_pmat(C, nroe, None, 0.0)
    ring_F = [np.zeros((N, N)) for _ in range(diis_size)]
    ring_e = [np.zeros((N, N)) for _ in range(diis_size)]
    damp, diis, it = damp0, False, 0
    DE, Eold = 10.0, 100.0
    energies, errors, coeffs = [], [], None
    eps = None
    while it < maxiter and abs(DE) > e_tol:
        P = build_pmat(C, nroe, P, damp)
        F = fock(P)
        ring_F = ring_F[1:] + ring_F[:1]
        ring_e = ring_e[1:] + ring_e[:1]
        ring_F[-1] = F.copy()
        ring_e[-1] = calc_diis_error(F, P, S)
        error = float(np.sum(np.abs(ring_e[-1]))) / (N * N)
        if error < diis_thresh and it >= diis_size:
            if not diis:
                diis, damp = True, 0.0
            m = diis_size
            B = np.zeros((m + 1, m + 1))
            for i in range(m):
                for j in range(m):
                    B[i, j] = float(np.dot(ring_e[i].reshape(-1), ring_e[j].reshape(-1)))
            B[:m, m] = B[m, :m] = -1.0
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            coeffs = np.linalg.solve(B, rhs)[:m]
            F = sum(c * Fj for c, Fj in zip(coeffs, ring_F))
        Escf = calc_e_el(F, P, H)
        DE, Eold = Escf - Eold, Escf
        energies.append(Escf + ion_rep)
        errors.append(error)
        C, eps = diag_fmat(F, Som12)
        it += 1
    P = build_pmat(C, nroe, P, damp)
    F = fock(P)
    one = float(np.sum(P * H))
    two = float(np.sum(P * F))
    return {"C": C, "eps": eps, "P": P, "F": F, "energies": np.array(energies), "errors": np.array(errors),
            "iterations": it, "diis": diis, "last_coeffs": coeffs, "one": one + ion_rep, "two": (two - one) / 2,
            "total": (one + two) / 2 + ion_rep, "converged": abs(DE) <= e_tol}
