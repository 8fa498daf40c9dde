"""numpy FP64 restatement of the reference's post-HF arithmetic (test oracle).

Each function names the reference lines (Masterluke87/corr) it follows.
Pins (checks that do not retype the formula; see tests/test_oracle.py):
  build_fmo          C = I gives F; unitary invariance of the trace.
  sorted_unique_list / unfold_sorted
                     round trip dense -> list -> dense on random 8-fold
                     symmetric tensors; class sizes against closed-form counts.
  motrans            brute force calc_mo2int_bf on N <= 6; C = I identity;
                     8-fold symmetry of the output; two orders of contraction.
  canonical_mp2      closed form on a 1-occupied/1-virtual system;
                     RI == canonical when (ia|jb) := sum_Q B B exactly.
  ri_metric          J^-1/2 J J^-1/2 = I; eigen vs Cholesky give the same B^T B;
                     the 1e-10 cutoff on a rank-deficient J.
  transform_ri       explicit quadruple loop on tiny sizes.
  rimp2              per-pair loop of the reference vs vectorised einsum;
                     RI == canonical.
  rimp2_gradient     central finite differences of rimp2.
"""
from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------- a1 ----------


def build_fmo(C: np.ndarray, F: np.ndarray) -> np.ndarray:
    """FMo[i,j] = sum_{mu nu} C[i,mu] F[mu,nu] C[j,nu]  (motrans.cc:8-32)."""
    return C @ F @ C.T


# ------------------------------------------------------ sorted unique list ----


def sorted_unique_list(dense: np.ndarray, thresh: float = 1.0e-12):
    """Sorted unique AO ERI list in the reference's format.

    Entry selection follows calculate_libint_tei (ops_rhf.cc:682-693): indices
    (a,b,c,d) with a >= b, c >= d, ab >= cd (ab = a(a+1)/2 + b) and
    |value| > thresh.  The reordering and the partition into the five
    permutation classes follow resort_integrals (ops_rhf.cc:760-896):
      class 0: a==b==c==d; 1: a==b and c==d; 2: a==c and b==d; 3: a==b; 4: rest.
    Returns (intval float64[n], intnums uint16[n,4], sortcount int64[4]).
    """
    N = dense.shape[0]
    a, b, c, d = np.meshgrid(*(np.arange(N),) * 4, indexing="ij")
    a, b, c, d = (x.ravel() for x in (a, b, c, d))
    ab = a * (a + 1) // 2 + b
    cd = c * (c + 1) // 2 + d
    keep = (a >= b) & (c >= d) & (ab >= cd)
    a, b, c, d = a[keep], b[keep], c[keep], d[keep]
    val = dense[a, b, c, d]
    nz = np.abs(val) > thresh
    a, b, c, d, val = a[nz], b[nz], c[nz], d[nz], val[nz]
    idx = np.stack([a, b, c, d], axis=1).astype(np.int64)
    idx = _resort_step1(idx)
    a, b, c, d = idx.T
    cls = np.full(len(val), 4, dtype=np.int64)
    # sequential partition: a test only sees entries not claimed earlier
    m0 = (a == b) & (b == c) & (c == d)
    cls[m0] = 0
    m1 = ~m0 & (a == b) & (c == d)
    cls[m1] = 1
    m2 = ~m0 & ~m1 & (a == c) & (b == d)
    cls[m2] = 2
    m3 = ~m0 & ~m1 & ~m2 & (a == b)
    cls[m3] = 3
    order = np.argsort(cls, kind="stable")
    sortcount = np.array([np.sum(cls <= k) for k in range(4)], dtype=np.int64)
    return val[order].astype(np.float64), idx[order].astype(np.uint16), sortcount


def _resort_step1(idx: np.ndarray) -> np.ndarray:
    """Index reorderings of resort_integrals step 1 (ops_rhf.cc:766-818).

    Every test reads the original (a,b,c,d); later assignments win."""
    a, b, c, d = (idx[:, k].copy() for k in range(4))
    out = idx.copy()

    def assign(mask, p, q, r, s):
        out[mask, 0], out[mask, 1], out[mask, 2], out[mask, 3] = p[mask], q[mask], r[mask], s[mask]

    assign((a == b) & (b == d) & (c != d), a, a, a, c)            # IIb -> IIa
    assign((a != b) & (a == c) & (c == d), a, a, a, b)            # IIc -> IIa
    assign((a != b) & (b == c) & (c == d), b, b, b, a)            # IId -> IIa
    assign((a == d) & (b == c) & (a != b), a, b, d, c)            # IIIc -> IIIb
    assign((b == c) & (b != a) & (b != d) & (a != d), b, a, c, d)  # IVc -> IVb
    assign((c == d) & (c != a) & (c != b) & (a != b), c, c, a, b)  # IVd -> IVa
    assign((a == d) & (a != b) & (a != c) & (b != c), a, b, d, c)  # IVe -> IVb
    assign((b == d) & (b != a) & (b != c) & (a != c), b, a, d, c)  # IVf -> IVb
    return out


# images of one entry per class, as (p q | r s) index permutations of (a,b,c,d)
# (trans_I, motrans.cc:358-402)
_IMAGES = {
    0: [(0, 1, 2, 3)],
    1: [(0, 1, 2, 3), (2, 3, 0, 1)],
    2: [(0, 1, 2, 3), (0, 1, 3, 2), (1, 0, 2, 3), (1, 0, 3, 2)],
    3: [(0, 1, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 0, 1)],
    4: [(0, 1, 2, 3), (0, 1, 3, 2), (1, 0, 2, 3), (1, 0, 3, 2),
        (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)],
}


def entry_class(sortcount: np.ndarray, n: int) -> np.ndarray:
    x = np.arange(n)
    return np.searchsorted(np.asarray(sortcount), x, side="right")


def unfold_sorted(N: int, intval, intnums, sortcount) -> np.ndarray:
    """Dense (ab|cd) from the sorted unique list, accumulating every image of
    every entry as trans_I does with C = identity (motrans.cc:341-403)."""
    dense = np.zeros((N, N, N, N))
    idx = np.asarray(intnums, dtype=np.int64).reshape(-1, 4)
    cls = entry_class(sortcount, len(intval))
    for k, perms in _IMAGES.items():
        m = cls == k
        for p in perms:
            np.add.at(dense, (idx[m, p[0]], idx[m, p[1]], idx[m, p[2]], idx[m, p[3]]), intval[m])
    return dense


# ------------------------------------------------------------- a2 - a5 ----


def motrans(C: np.ndarray, dense: np.ndarray) -> np.ndarray:
    """prec_ints[i,j,k,l] = sum C_ia C_jb C_kc C_ld (ab|cd), contracted one index
    at a time in the order of mo2int_op (motrans.cc:407-503): a, then b, c, d."""
    t = np.tensordot(C, dense, axes=([1], [0]))          # (i|bcd)
    t = np.tensordot(C, t, axes=([1], [1]))              # (j, i, c, d)
    t = t.transpose(1, 0, 2, 3)                           # (ij|cd)
    t = np.tensordot(t, C, axes=([2], [1]))              # (i, j, d, k)
    t = t.transpose(0, 1, 3, 2)                           # (ij|kd)
    return np.tensordot(t, C, axes=([3], [1]))            # (ij|kl)


def calc_mo2int_bf(i, j, k, l, C, intval, intnums, sortcount) -> float:
    """Brute-force single MO integral over the unique list (motrans.cc:225-327)."""
    idx = np.asarray(intnums, dtype=np.int64).reshape(-1, 4)
    cls = entry_class(sortcount, len(intval))
    total = 0.0
    for k_, perms in _IMAGES.items():
        m = cls == k_
        for p in perms:
            total += np.sum(C[i, idx[m, p[0]]] * C[j, idx[m, p[1]]] * C[k, idx[m, p[2]]] *
                            C[l, idx[m, p[3]]] * intval[m])
    return float(total)


# --------------------------------------------------------------- a6 ----------


def mp2_from_iajb(iajb: np.ndarray, eo: np.ndarray, ev: np.ndarray) -> dict:
    """Energies from (ia|jb)[i,a,j,b] (can_mp2.cc:26-46 arithmetic)."""
    D = ev[None, :, None, None] + ev[None, None, None, :] - eo[:, None, None, None] - eo[None, None, :, None]
    ibja = iajb.transpose(0, 3, 2, 1)
    e_os = -np.sum(iajb * iajb / D)
    e_ss = -np.sum((iajb - ibja) * iajb / D)
    return _energies(e_os, e_ss)


def canonical_mp2(prec_ints: np.ndarray, FMo: np.ndarray, nocc: int) -> dict:
    """run_canonical_mp2 (can_mp2.cc:11-49): (ia|jb) read from prec_ints at
    [i, o+a, j, o+b], denominators from the FMo diagonal."""
    o = nocc
    e = np.diag(FMo)
    iajb = prec_ints[:o, o:, :o, o:]
    return mp2_from_iajb(iajb, e[:o], e[o:])


def _energies(e_os: float, e_ss: float) -> dict:
    return {"EMP2_OS": float(e_os), "EMP2_SS": float(e_ss), "EMP2": float(e_os + e_ss),
            "EMP2_SCS": float(e_ss / 3.0 + e_os * 6.0 / 5.0)}


# --------------------------------------------------------------- a7 ----------


def ri_metric_eigen(J: np.ndarray, tol: float = 1.0e-10):
    """J^-1/2 = V diag(s) V^T with s = 0 if l/l_max < tol or l < 0, else l^-1/2
    (ri.cc:50-69).  Returns (M, number of kept eigenvalues)."""
    lam, V = np.linalg.eigh(J)
    lmax = lam[-1]
    keep = ~((lam / lmax < tol) | (lam < 0))
    s = np.zeros_like(lam)
    s[keep] = 1.0 / np.sqrt(lam[keep])
    return (V * s) @ V.T, int(keep.sum())


def ri_metric_cholesky(J: np.ndarray) -> np.ndarray:
    """M = L^-1 with J = L L^T, so that M^T M = J^-1."""
    L = np.linalg.cholesky(J)
    return np.linalg.inv(L)


def ri_bpq(M: np.ndarray, Pmn: np.ndarray) -> np.ndarray:
    """BPQ = J^-1/2 (P|mu nu)  (ri.cc:102-103); Pmn is [naux, N, N]."""
    naux, N, _ = Pmn.shape
    return (M @ Pmn.reshape(naux, N * N)).reshape(naux, N, N)


# --------------------------------------------------------------- a8 ----------


def transform_ri(C: np.ndarray, BPQ: np.ndarray, nocc: int) -> np.ndarray:
    """Bia[Q,i,a] = sum_{mu nu} C[i,mu] C[o+a,nu] BPQ[Q,mu,nu] via BiQ
    (ri.cc:147-172).  Returned layout [Q, i, a]."""
    o = nocc
    BiQ = np.einsum("im,Qmn->Qin", C[:o], BPQ, optimize=True)
    return np.einsum("an,Qin->Qia", C[o:], BiQ, optimize=True)


# --------------------------------------------------------------- a9 ----------


def rimp2(Bia: np.ndarray, eo: np.ndarray, ev: np.ndarray, pairs=None) -> dict:
    """run_canonical_mp2_ri (can_mp2.cc:53-132): for every ordered pair (i,j)
    one GEMM M = B_i^T B_j over Q, X = 1/D, MD = M - M^T, then
    EMP2_OS -= <X o M, M> and EMP2_SS -= <X o M, MD>.

    ``pairs`` restricts the loop to a subset of ordered pairs (used to time a
    bounded sample of the CPU path); the energies are then partial sums."""
    o = Bia.shape[1]
    e_os = 0.0
    e_ss = 0.0
    if pairs is None:
        pairs = [(i, j) for i in range(o) for j in range(o)]
    for i, j in pairs:
        M = Bia[:, i, :].T @ Bia[:, j, :]
        X = 1.0 / (ev[:, None] + ev[None, :] - eo[i] - eo[j])
        MD = M - M.T
        tmp = X * M
        e_os -= np.dot(tmp.ravel(), M.ravel())
        e_ss -= np.dot(tmp.ravel(), MD.ravel())
    return _energies(e_os, e_ss)


def rimp2_dense(Bia: np.ndarray, eo: np.ndarray, ev: np.ndarray) -> dict:
    """Same energies through the full (ia|jb) tensor (small sizes only)."""
    iajb = np.einsum("Qia,Qjb->iajb", Bia, Bia, optimize=True)
    return mp2_from_iajb(iajb, eo, ev)


def rimp2_gradient(Bia: np.ndarray, eo: np.ndarray, ev: np.ndarray, c_os: float = 1.0,
                   c_ss: float = 1.0):
    """Analytic gradient of E = c_os EMP2_OS + c_ss EMP2_SS of rimp2.

    With M^{ij}_ab = sum_Q B^Q_ia B^Q_jb, D = e_a + e_b - e_i - e_j,
      R^{ij}_ab = [(c_os + c_ss) M_ab - c_ss M_ba] / D
      N^{ij}_ab = (c_os + c_ss) M_ab^2 - c_ss M_ab M_ba
      dE/dB^Q_ia = -4 sum_{jb} R^{ij}_ab B^Q_jb
      dE/de_a    =  2 sum_{ijb} N^{ij}_ab / D^2
      dE/de_i    = -2 sum_{jab} N^{ij}_ab / D^2
    Returns (E, dB, d_eo, d_ev)."""
    M = np.einsum("Qia,Qjb->ijab", Bia, Bia, optimize=True)
    D = ev[None, None, :, None] + ev[None, None, None, :] - eo[:, None, None, None] - eo[None, :, None, None]
    Mt = M.transpose(0, 1, 3, 2)
    R = ((c_os + c_ss) * M - c_ss * Mt) / D
    Nn = (c_os + c_ss) * M * M - c_ss * M * Mt
    E = -np.sum(Nn / D)
    dB = -4.0 * np.einsum("ijab,Qjb->Qia", R, Bia, optimize=True)
    W = Nn / (D * D)
    d_ev = 2.0 * np.einsum("ijab->a", W)
    d_eo = -2.0 * np.einsum("ijab->i", W)
    return float(E), dB, d_eo, d_ev


# ------------------------------------------------------ non-canonical MP2 ----


def non_canonical_mp2(prec: np.ndarray, FMo: np.ndarray, nocc: int, iters: int = 20):
    """Semi-canonical MP2 amplitudes with off-diagonal occupied Fock coupling
    (run_non_canonical_mp2, loc_mp2.cc:7-92).  Returns (E[iters], Rsum[iters],
    T) where iteration n prints E and Rsum of the amplitudes before its
    update:
      T0 = -(ia|jb) / D,  D = f_aa + f_bb - f_ii - f_jj
      G_ijab = -sum_{k != i} f_ik T_kjab
      R = (ia|jb) + D T + G_ijab + G_jiba,  Rsum = sum |R|
      E = sum ((ia|jb) + R)(2 T_ijab - T_ijba),  T -= R / D."""
    N = FMo.shape[0]
    o, v = nocc, N - nocc
    K = prec[:o, o:, :o, o:].transpose(0, 2, 1, 3).copy()        # [i][j][a][b] = (ia|jb)
    f = np.diag(FMo)
    D = f[o:][None, None, :, None] + f[o:][None, None, None, :] - f[:o][:, None, None, None] \
        - f[:o][None, :, None, None]
    T = -K / D
    Foff = FMo[:o, :o].copy()
    np.fill_diagonal(Foff, 0.0)
    E, Rs = [], []
    for _ in range(iters):
        G = -np.einsum("ik,kjab->ijab", Foff, T)
        R = K + D * T + G + G.transpose(1, 0, 3, 2)
        Rs.append(float(np.sum(np.abs(R))))
        E.append(float(np.sum((K + R) * (2 * T - T.transpose(0, 1, 3, 2)))))
        T = T - R / D
    return np.array(E), np.array(Rs), T
