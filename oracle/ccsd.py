"""numpy FP64 restatement of the reference's spin-orbital CCSD (test oracle).

Follows src/POSTHF/ccsd.cc:6-580 and src/POSTHF/ccsd_intermediates.cc:6-287
(Stanton & Gauss, J. Chem. Phys. 94, 4334 (1991)) term by term, including the
Fae quirk of ccsd_intermediates.cc:33 (f_{m a} t_m^a instead of f_{m e} t_m^a),
the T2 update that reads the old T1, and the energy evaluated at the top of
each iteration.  Spin orbital p is spatial p//2 with spin p%2; the nocc = nroe
lowest spin orbitals are occupied.

Pins (tests/test_oracle_ccsd.py): the guess energy equals spatial MP2 of
oracle.posthf.canonical_mp2; for two electrons converged CCSD equals full CI
(oracle.fci); T2 stays antisymmetric; the converged residual vanishes.
"""
from __future__ import annotations

import numpy as np


def ints_so(prec_ints: np.ndarray) -> np.ndarray:
    """<pq||rs> = (pr|qs) d(sp,sr) d(sq,ss) - (ps|qr) d(sp,ss) d(sq,sr)
    over n = 2N spin orbitals (get_integral / calc_ints_so, ccsd.cc:6-9, 509-535)."""
    N = prec_ints.shape[0]
    n = 2 * N
    sp = np.arange(n) // 2
    sg = np.arange(n) % 2
    chem = prec_ints[np.ix_(sp, sp, sp, sp)]                 # (p r | q s) indexed [p, r, q, s]
    same = (sg[:, None] == sg[None, :]).astype(float)
    coul = chem.transpose(0, 2, 1, 3) * same[:, None, :, None] * same[None, :, None, :]
    # coul[p,q,r,s] = (pr|qs) d(p,r) d(q,s)
    return coul - coul.transpose(0, 1, 3, 2)


def form_fock(C: np.ndarray, H: np.ndarray, V: np.ndarray, nocc: int):
    """Spin-orbital h and f = h + sum_{j<nocc} <pj||qj> (ccsd.cc:11-48)."""
    hs = C @ H @ C.T
    n = 2 * C.shape[0]
    sp = np.arange(n) // 2
    sg = np.arange(n) % 2
    h = hs[np.ix_(sp, sp)] * (sg[:, None] == sg[None, :])
    f = h + np.einsum("pjqj->pq", V[:, :nocc, :, :nocc])
    return h, f


def e_hf_check(prec_ints: np.ndarray, h: np.ndarray, nocc: int, ion_rep: float) -> float:
    """El of calc_e_check (ccsd.cc:50-81)."""
    i = np.arange(nocc)
    si, sg = i // 2, i % 2
    J = prec_ints[si[:, None], si[:, None], si[None, :], si[None, :]]
    K = prec_ints[si[:, None], si[None, :], si[:, None], si[None, :]] * (sg[:, None] == sg[None, :])
    return float(np.trace(h[:nocc, :nocc]) + 0.5 * np.sum(J - K) + ion_rep)


class Blocks:
    """Views of <pq||rs> and f split into occupied (o) / virtual (v) ranges."""

    def __init__(self, V: np.ndarray, f: np.ndarray, nocc: int):
        o = nocc
        O, W = slice(0, o), slice(o, V.shape[0])
        self.o, self.v = o, V.shape[0] - o
        self.oooo = V[O, O, O, O]
        self.ooov = V[O, O, O, W]
        self.oovo = V[O, O, W, O]
        self.oovv = V[O, O, W, W]
        self.ovoo = V[O, W, O, O]
        self.ovov = V[O, W, O, W]
        self.ovvo = V[O, W, W, O]
        self.ovvv = V[O, W, W, W]
        self.vovv = V[W, O, W, W]
        self.vvvo = V[W, W, W, O]
        self.vvvv = V[W, W, W, W]
        self.foo, self.fov, self.fvv = f[O, O], f[O, W], f[W, W]
        fd = np.diag(f)
        self.eo, self.ev = fd[:o], fd[o:]


def denominators(B: Blocks):
    d1 = B.eo[:, None] - B.ev[None, :]
    d2 = B.eo[:, None, None, None] + B.eo[None, :, None, None] - B.ev[None, None, :, None] - B.ev[None, None, None, :]
    return d1, d2


def guess(B: Blocks):
    """T1 = 0, T2 = <ij||ab> / D_ijab and the spin-orbital MP2 energy
    1/4 sum <ij||ab> T2 (ccsd_guess, ccsd.cc:84-144)."""
    _, d2 = denominators(B)
    T2 = B.oovv / d2
    return np.zeros((B.o, B.v)), T2, 0.25 * float(np.sum(B.oovv * T2))


def energy(B: Blocks, T1, T2) -> float:
    """ccsd_energy (ccsd.cc:147-219)."""
    e = np.sum(B.fov * T1)
    e += 0.25 * np.sum(T2 * B.oovv)
    e += 0.5 * np.einsum("lc,kd,kldc->", T1, T1, B.oovv)
    return float(e)


def taus(T1, T2):
    """tau~ and tau (ccsd_build_taus, ccsd.cc:221-251)."""
    x = np.einsum("ia,jb->ijab", T1, T1)
    xx = x - x.transpose(0, 1, 3, 2)
    return T2 + 0.5 * xx, T2 + xx


def intermediates(B: Blocks, T1, T2, tau_s, tau):
    """Fae, Fmi, Fme, Wmnij, Wabef, Wmbej (ccsd_intermediates.cc:6-287)."""
    v, o = B.v, B.o
    Fae = B.fvv * (1.0 - np.eye(v))
    Fae -= 0.5 * np.einsum("ma,ma->a", B.fov, T1)[:, None]          # quirk of :33
    Fae += np.einsum("mf,mafe->ae", T1, B.ovvv)
    Fae -= 0.5 * np.einsum("mnaf,mnef->ae", tau_s, B.oovv)

    Fmi = B.foo * (1.0 - np.eye(o))
    Fmi += 0.5 * np.einsum("ie,me->mi", T1, B.fov)
    Fmi += np.einsum("ne,mnie->mi", T1, B.ooov)
    Fmi += 0.5 * np.einsum("inef,mnef->mi", tau_s, B.oovv)

    Fme = B.fov + np.einsum("nf,mnef->me", T1, B.oovv)

    Wmnij = B.oooo.copy()
    x = np.einsum("je,mnie->mnij", T1, B.ooov)
    Wmnij += x - x.transpose(0, 1, 3, 2)
    Wmnij += 0.25 * np.einsum("ijef,mnef->mnij", tau, B.oovv)

    y = np.einsum("mb,amef->abef", T1, B.vovv)
    Wabef = B.vvvv - y + y.transpose(1, 0, 2, 3)
    Wabef += 0.25 * np.einsum("mnab,mnef->abef", tau, B.oovv)

    Wmbej = B.ovvo.copy()
    Wmbej += np.einsum("jf,mbef->mbej", T1, B.ovvv)
    Wmbej -= np.einsum("nb,mnej->mbej", T1, B.oovo)
    Z = 0.5 * T2 + np.einsum("jf,nb->jnfb", T1, T1)
    Wmbej -= np.einsum("jnfb,mnef->mbej", Z, B.oovv)
    return Fae, Fmi, Fme, Wmnij, Wabef, Wmbej


def update(B: Blocks, T1, T2):
    """One pass of taus, intermediates, ccsd_update_T1 and ccsd_update_T2
    (ccsd.cc:221-467).  Returns (T1n, T2n)."""
    tau_s, tau = taus(T1, T2)
    Fae, Fmi, Fme, Wmnij, Wabef, Wmbej = intermediates(B, T1, T2, tau_s, tau)
    d1, d2 = denominators(B)

    T1n = B.fov.copy()
    T1n += np.einsum("ie,ae->ia", T1, Fae)
    T1n -= np.einsum("ma,mi->ia", T1, Fmi)
    T1n += np.einsum("imae,me->ia", T2, Fme)
    T1n -= np.einsum("nf,naif->ia", T1, B.ovov)
    T1n -= 0.5 * np.einsum("imef,maef->ia", T2, B.ovvv)
    T1n -= 0.5 * np.einsum("mnae,nmei->ia", T2, B.oovo)
    T1n /= d1

    Faep = Fae - 0.5 * np.einsum("mb,me->be", T1, Fme)
    Fmip = Fmi + 0.5 * np.einsum("je,me->mj", T1, Fme)
    T2n = B.oovv.copy()
    x = np.einsum("ijae,be->ijab", T2, Faep)
    T2n += x - x.transpose(0, 1, 3, 2)
    x = np.einsum("imab,mj->ijab", T2, Fmip)
    T2n -= x - x.transpose(1, 0, 2, 3)
    T2n += 0.5 * np.einsum("mnab,mnij->ijab", tau, Wmnij)
    T2n += 0.5 * np.einsum("ijef,abef->ijab", tau, Wabef)
    Z = np.einsum("imae,mbej->iajb", T2, Wmbej)
    Z -= np.einsum("ie,ma,mbej->iajb", T1, T1, B.ovvo)
    T2n += (Z.transpose(0, 2, 1, 3) - Z.transpose(0, 2, 3, 1)
            - Z.transpose(2, 0, 1, 3) + Z.transpose(2, 0, 3, 1))
    x = np.einsum("ie,abej->ijab", T1, B.vvvo)
    T2n += x - x.transpose(1, 0, 2, 3)
    x = np.einsum("ma,mbij->ijab", T1, B.ovoo)
    T2n -= x - x.transpose(0, 1, 3, 2)
    T2n /= d2
    return T1n, T2n


def ccsd_ur(prec_ints, C, H, nroe, ion_rep, maxiter=20):
    """ccsd_ur (ccsd.cc:538-580): guess, then maxiter Jacobi updates with the
    energy of the current amplitudes evaluated at the top of each iteration.
    Returns dict with the per-iteration energies, MP2, the HF check and the
    final amplitudes (after maxiter updates)."""
    V = ints_so(prec_ints)
    h, f = form_fock(C, H, V, nroe)
    B = Blocks(V, f, nroe)
    T1, T2, emp2 = guess(B)
    energies = []
    for _ in range(maxiter):
        energies.append(energy(B, T1, T2))
        T1, T2 = update(B, T1, T2)
    return {"energies": energies, "emp2": emp2, "e_hf_check": e_hf_check(prec_ints, h, nroe, ion_rep),
            "T1": T1, "T2": T2, "f": f, "V": V}


def ccsd_converge(B: Blocks, maxiter=200, tol=1e-12, diis=8):
    """Converged CCSD by Jacobi updates with DIIS on (T1, T2) (test reference
    for the DIIS mode; the reference itself has no convergence loop)."""
    T1, T2, _ = guess(B)
    vecs, errs = [], []
    e_old = energy(B, T1, T2)
    for it in range(maxiter):
        T1n, T2n = update(B, T1, T2)
        r = np.concatenate([(T1n - T1).ravel(), (T2n - T2).ravel()])
        vecs.append(np.concatenate([T1n.ravel(), T2n.ravel()]))
        errs.append(r)
        if len(vecs) > diis:
            vecs.pop(0)
            errs.pop(0)
        if len(vecs) >= 2:
            m = len(vecs)
            Bm = -np.ones((m + 1, m + 1))
            Bm[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    Bm[a, b] = errs[a] @ errs[b]
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            c = np.linalg.solve(Bm, rhs)[:m]
            x = sum(ci * vi for ci, vi in zip(c, vecs))
        else:
            x = vecs[-1]
        T1 = x[:T1.size].reshape(T1.shape)
        T2 = x[T1.size:].reshape(T2.shape)
        e = energy(B, T1, T2)
        if abs(e - e_old) < tol and np.linalg.norm(r) < 1e-9:
            return e, T1, T2, it + 1
        e_old = e
    return e_old, T1, T2, maxiter
