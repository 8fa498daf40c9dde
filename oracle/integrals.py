"""numpy restatement of the Gaussian integrals the reference takes from libint2
(test oracle: only tests/, __graft_entry__.smoke() and bench.py's CPU leg use it).

The reference evaluates with libint2 2.4.2 (not vendored): one-electron S, T,
V (calculate_libint_oei, src/SCF/ops_rhf.cc:77-103), four-centre (mu nu|la si)
(calculate_libint_tei, ops_rhf.cc:618-723) and the RI integrals (P|Q) and
(P|mu nu) (calculate_ri, src/POSTHF/ri.cc:25-100), over spherical ("pure",
main.cc:63) contracted Gaussians read from basissets/*.g94.  This file
restates the published McMurchie-Davidson scheme (McMurchie & Davidson,
J. Comput. Phys. 26, 218 (1978); Helgaker, Joergensen & Olsen, Molecular
Electronic-Structure Theory, ch. 9): Hermite expansion coefficients E^{ij}_t,
Hermite Coulomb integrals R_tuv from the Boys function, real solid harmonics
by the recursion of Helgaker eq. 6.4.70-6.4.72, every contracted spherical
function normalised to unit self-overlap (libint2's convention).

Pins (tests/test_oracle_integrals.py):
  STO-3G H2 at R = 1.4 bohr against the textbook values of Szabo & Ostlund
  (Modern Quantum Chemistry, 1989, eqs. 3.229-3.235 and E = -1.1167 Eh);
  the raw self-overlap of every solid-harmonic primitive (l <= 4) against its
  closed form 4 pi/(2l+1) Int r^(2l+2) e^(-2 a r^2) dr;
  same-centre Coulomb integrals (l <= 3) against radial quadrature of the
  multipole expansion;  rotational and translational invariance of the RHF
  and RI-MP2 energies of H2O in def2-SVP / def2-SVP-RI (d and f shells);
  8-fold symmetry of (mu nu|la si), (P|Q) = (Q|P), RI == exact integrals for
  products of one aux function.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
from scipy.special import gammainc, gamma as gamma_fn

ANGSTROM_TO_BOHR = 1.0 / 0.52917721067   # CODATA 2014, as psi4 1.1-1.3 used
_L = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}
Z_OF = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


# ---------------------------------------------------------------- basis ----

def parse_g94(path: str) -> dict:
    """{element: [(l, exps, coefs), ...]} from a Gaussian-94 basis file.
    An "SP" shell yields an s and a p shell with the same exponents."""
    out, el, lines = {}, None, open(path).read().splitlines()
    i = 0
    while i < len(lines):
        t = lines[i].split()
        i += 1
        if not t or t[0].startswith("!"):
            continue
        if t[0] == "****":
            el = None
            continue
        if el is None:
            if t[0] in Z_OF and len(t) >= 2 and t[1] == "0":
                el = t[0]
                out[el] = []
            continue
        kind, nprim = t[0].upper(), int(t[1])
        rows = []
        for _ in range(nprim):
            rows.append([float(x.replace("D", "E").replace("d", "e")) for x in lines[i].split()])
            i += 1
        rows = np.array(rows)
        if kind == "SP":
            out[el].append((0, rows[:, 0].copy(), rows[:, 1].copy()))
            out[el].append((1, rows[:, 0].copy(), rows[:, 2].copy()))
        else:
            out[el].append((_L[kind], rows[:, 0].copy(), rows[:, 1].copy()))
    return out


@dataclass
class Shell:
    l: int
    center: np.ndarray      # bohr
    exps: np.ndarray
    coefs: np.ndarray       # include the primitive normalisation of x^l e^{-a r^2}
    atom: int


def _dfact(n: int) -> float:
    return float(np.prod(np.arange(n, 0, -2))) if n > 0 else 1.0


def build_shells(atoms, basis: dict) -> list:
    """atoms: [(symbol, xyz_bohr)]; shells in atom order, then file order."""
    shells = []
    for a, (sym, xyz) in enumerate(atoms):
        for l, e, c in basis[sym]:
            norm = (2 * e / np.pi) ** 0.75 * (4 * e) ** (l / 2) / math.sqrt(_dfact(2 * l - 1))
            shells.append(Shell(l, np.asarray(xyz, float), np.asarray(e, float), np.asarray(c * norm, float), a))
    return shells


def nbf(shells) -> int:
    return sum(2 * s.l + 1 for s in shells)


def offsets(shells):
    o, out = 0, []
    for s in shells:
        out.append(o)
        o += 2 * s.l + 1
    return out


def cart_list(l: int):
    """Cartesian exponents (i, j, k), i + j + k = l, x-major order."""
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


@lru_cache(maxsize=None)
def solid_harmonics(l: int) -> np.ndarray:
    """(2l+1) x ncart matrix: row m + l holds the Cartesian coefficients of the
    real regular solid harmonic S_lm (Helgaker eq. 6.4.70-6.4.72)."""
    def mul(poly, ijk, f):
        out = {}
        for k, v in poly.items():
            kk = (k[0] + ijk[0], k[1] + ijk[1], k[2] + ijk[2])
            out[kk] = out.get(kk, 0.0) + v * f
        return out

    def add(*ps):
        out = {}
        for p in ps:
            for k, v in p.items():
                out[k] = out.get(k, 0.0) + v
        return out

    S = {(0, 0): {(0, 0, 0): 1.0}}
    for ll in range(0, l):
        d = 1.0 if ll == 0 else 0.0
        f = math.sqrt(2.0 ** d * (2 * ll + 1) / (2 * ll + 2))
        S[(ll + 1, ll + 1)] = add(mul(S[(ll, ll)], (1, 0, 0), f), mul(S[(ll, -ll)], (0, 1, 0), -f * (1 - d)))
        S[(ll + 1, -ll - 1)] = add(mul(S[(ll, ll)], (0, 1, 0), f), mul(S[(ll, -ll)], (1, 0, 0), f * (1 - d)))
        for m in range(-ll, ll + 1):
            den = math.sqrt((ll + m + 1) * (ll - m + 1))
            t1 = mul(S[(ll, m)], (0, 0, 1), (2 * ll + 1) / den)
            if ll >= 1 and abs(m) <= ll - 1:
                g = -math.sqrt((ll + m) * (ll - m)) / den
                r2 = add(mul(S[(ll - 1, m)], (2, 0, 0), g), mul(S[(ll - 1, m)], (0, 2, 0), g),
                         mul(S[(ll - 1, m)], (0, 0, 2), g))
                S[(ll + 1, m)] = add(t1, r2)
            else:
                S[(ll + 1, m)] = t1
    cl = cart_list(l)
    T = np.zeros((2 * l + 1, len(cl)))
    for m in range(-l, l + 1):
        for k, v in S[(l, m)].items():
            T[m + l, cl.index(k)] += v
    return T


# ---------------------------------------------------------- MD primitives ----

def boys(nmax: int, T: float) -> np.ndarray:
    """F_n(T), n = 0..nmax, from the regularised incomplete gamma function."""
    n = np.arange(nmax + 1, dtype=float)
    if T < 1e-12:
        return 1.0 / (2 * n + 1) - T / (2 * n + 3)
    return gammainc(n + 0.5, T) * gamma_fn(n + 0.5) / (2.0 * T ** (n + 0.5))


def hermite_E(la: int, lb: int, a: float, b: float, A: float, B: float) -> np.ndarray:
    """E[i, j, t] of x_A^i x_B^j e^{-a x_A^2 - b x_B^2} = sum_t E_t Lambda_t (1D)."""
    p = a + b
    P = (a * A + b * B) / p
    XPA, XPB = P - A, P - B
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = math.exp(-a * b / p * (A - B) ** 2)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t] + (t + 1) * E[i - 1, j, t + 1]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                else:
                    v = XPB * E[i, j - 1, t] + (t + 1) * E[i, j - 1, t + 1]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                E[i, j, t] = v
    return E


def hermite_R(L: int, alpha: float, X: float, Y: float, Z: float) -> np.ndarray:
    """R[t, u, v] = R^0_tuv(alpha, (X,Y,Z)) for t + u + v <= L."""
    F = boys(L, alpha * (X * X + Y * Y + Z * Z))
    R = np.zeros((L + 1, L + 1, L + 1, L + 1))   # [n, t, u, v]
    for n in range(L + 1):
        R[n, 0, 0, 0] = (-2 * alpha) ** n * F[n]
    for s in range(1, L + 1):                       # total order t + u + v = s
        for n in range(L - s, -1, -1):
            for t in range(s + 1):
                for u in range(s - t + 1):
                    v = s - t - u
                    if t > 0:
                        r = X * R[n + 1, t - 1, u, v] + ((t - 1) * R[n + 1, t - 2, u, v] if t > 1 else 0.0)
                    elif u > 0:
                        r = Y * R[n + 1, t, u - 1, v] + ((u - 1) * R[n + 1, t, u - 2, v] if u > 1 else 0.0)
                    else:
                        r = Z * R[n + 1, t, u, v - 1] + ((v - 1) * R[n + 1, t, u, v - 2] if v > 1 else 0.0)
                    R[n, t, u, v] = r
    return R[0]


def _pair_E(sa: Shell, sb: Shell, ia: int, ib: int):
    a, b = sa.exps[ia], sb.exps[ib]
    Ex = hermite_E(sa.l, sb.l, a, b, sa.center[0], sb.center[0])
    Ey = hermite_E(sa.l, sb.l, a, b, sa.center[1], sb.center[1])
    Ez = hermite_E(sa.l, sb.l, a, b, sa.center[2], sb.center[2])
    return a + b, (a * sa.center + b * sb.center) / (a + b), Ex, Ey, Ez


def _pair_hermite(sa: Shell, sb: Shell, ia: int, ib: int):
    """(p, P, H) with H[ca, cb, t, u, v] the Hermite coefficients of the
    Cartesian product pair (ca, cb) of one primitive pair."""
    p, P, Ex, Ey, Ez = _pair_E(sa, sb, ia, ib)
    ca, cb = cart_list(sa.l), cart_list(sb.l)
    L = sa.l + sb.l
    H = np.zeros((len(ca), len(cb), L + 1, L + 1, L + 1))
    for x, (i, k, m) in enumerate(ca):
        for y, (j, l_, n) in enumerate(cb):
            H[x, y, :i + j + 1, :k + l_ + 1, :m + n + 1] = np.einsum(
                "t,u,v->tuv", Ex[i, j, :i + j + 1], Ey[k, l_, :k + l_ + 1], Ez[m, n, :m + n + 1])
    return p, P, H


# --------------------------------------------------------- shell blocks ----

def _sph(block: np.ndarray, ls) -> np.ndarray:
    for ax, l in enumerate(ls):
        block = np.moveaxis(np.tensordot(solid_harmonics(l), block, axes=([1], [ax])), 0, ax)
    return block


def overlap_block(sa, sb, raw=False):
    out = np.zeros((len(cart_list(sa.l)), len(cart_list(sb.l))))
    for ia in range(len(sa.exps)):
        for ib in range(len(sb.exps)):
            p, P, Ex, Ey, Ez = _pair_E(sa, sb, ia, ib)
            f = sa.coefs[ia] * sb.coefs[ib] * (np.pi / p) ** 1.5
            for x, (i, k, m) in enumerate(cart_list(sa.l)):
                for y, (j, l_, n) in enumerate(cart_list(sb.l)):
                    out[x, y] += f * Ex[i, j, 0] * Ey[k, l_, 0] * Ez[m, n, 0]
    return _sph(out, (sa.l, sb.l))


def kinetic_block(sa, sb):
    out = np.zeros((len(cart_list(sa.l)), len(cart_list(sb.l))))
    for ia in range(len(sa.exps)):
        for ib in range(len(sb.exps)):
            a, b = sa.exps[ia], sb.exps[ib]
            p = a + b
            Es = [hermite_E(sa.l, sb.l + 2, a, b, sa.center[d], sb.center[d]) for d in range(3)]
            f = sa.coefs[ia] * sb.coefs[ib] * (np.pi / p) ** 1.5
            for x, ci in enumerate(cart_list(sa.l)):
                for y, cj in enumerate(cart_list(sb.l)):
                    S1 = [Es[d][ci[d], cj[d], 0] for d in range(3)]
                    tot = 0.0
                    for d in range(3):
                        j = cj[d]
                        Td = 4 * b * b * Es[d][ci[d], j + 2, 0] - 2 * b * (2 * j + 1) * Es[d][ci[d], j, 0]
                        if j >= 2:
                            Td += j * (j - 1) * Es[d][ci[d], j - 2, 0]
                        prod = Td
                        for e in range(3):
                            if e != d:
                                prod *= S1[e]
                        tot += prod
                    out[x, y] += -0.5 * f * tot
    return _sph(out, (sa.l, sb.l))


def dipole_block(sa, sb, origin):
    """<a| r_d - C_d |b> for d = x, y, z: in Hermite form x_C = x_P + (P - C)
    gives E^{ij}_1 + X_PC E^{ij}_0 per dimension (times (pi/p)^(3/2)).
    Returns [3][cart a][cart b] before the spherical transform."""
    C = np.asarray(origin, float)
    out = np.zeros((3, len(cart_list(sa.l)), len(cart_list(sb.l))))
    for ia in range(len(sa.exps)):
        for ib in range(len(sb.exps)):
            a, b = sa.exps[ia], sb.exps[ib]
            p = a + b
            P = (a * sa.center + b * sb.center) / p
            Es = [hermite_E(sa.l, sb.l, a, b, sa.center[d], sb.center[d]) for d in range(3)]
            f = sa.coefs[ia] * sb.coefs[ib] * (np.pi / p) ** 1.5
            for x, ci in enumerate(cart_list(sa.l)):
                for y, cj in enumerate(cart_list(sb.l)):
                    S1 = [Es[d][ci[d], cj[d], 0] for d in range(3)]
                    for d in range(3):
                        m = Es[d][ci[d], cj[d], 1] + (P[d] - C[d]) * S1[d]
                        prod = m
                        for e in range(3):
                            if e != d:
                                prod *= S1[e]
                        out[d, x, y] += f * prod
    return np.stack([_sph(out[d], (sa.l, sb.l)) for d in range(3)])


def nuclear_block(sa, sb, charges, coords):
    out = np.zeros((len(cart_list(sa.l)), len(cart_list(sb.l))))
    L = sa.l + sb.l
    for ia in range(len(sa.exps)):
        for ib in range(len(sb.exps)):
            p, P, H = _pair_hermite(sa, sb, ia, ib)
            f = sa.coefs[ia] * sb.coefs[ib] * 2 * np.pi / p
            for Zc, C in zip(charges, coords):
                R = hermite_R(L, p, *(P - C))
                out += -Zc * f * np.einsum("xytuv,tuv->xy", H, R[:L + 1, :L + 1, :L + 1])
    return _sph(out, (sa.l, sb.l))


def _aux_hermite(s: Shell, ia: int):
    z = Shell(0, s.center, np.array([0.0]), np.array([1.0]), s.atom)
    g, _, H = _pair_hermite(s, z, ia, 0)
    return g, H[:, 0]          # [c, t, u, v]


def coulomb2_block(sP, sQ):
    """(P|Q) over two aux shells."""
    LP, LQ = sP.l, sQ.l
    out = np.zeros((len(cart_list(LP)), len(cart_list(LQ))))
    for ia in range(len(sP.exps)):
        g, HP = _aux_hermite(sP, ia)
        for ib in range(len(sQ.exps)):
            d, HQ = _aux_hermite(sQ, ib)
            alpha = g * d / (g + d)
            R = hermite_R(LP + LQ, alpha, *(sP.center - sQ.center))
            sign = _signs(LQ)
            Rs = _shift_loop(R, LP, LQ)
            f = sP.coefs[ia] * sQ.coefs[ib] * 2 * np.pi ** 2.5 / (g * d * math.sqrt(g + d))
            out += f * np.einsum("atuv,bTUV,tuvTUV->ab", HP[:, :LP + 1, :LP + 1, :LP + 1], HQ[:, :LQ + 1, :LQ + 1,
                                                                                          :LQ + 1] * sign, Rs)
    return _sph(out, (LP, LQ))


def _signs(L):
    t = np.arange(L + 1)
    return ((-1.0) ** (t[:, None, None] + t[None, :, None] + t[None, None, :]))


def _shift_loop(R, L1, L2):
    out = np.zeros((L1 + 1,) * 3 + (L2 + 1,) * 3)
    for t in range(L1 + 1):
        for u in range(L1 + 1):
            for v in range(L1 + 1):
                out[t, u, v] = R[t:t + L2 + 1, u:u + L2 + 1, v:v + L2 + 1]
    return out


def coulomb3_block(sP, sa, sb):
    """(P|ab) over an aux shell and an orbital shell pair."""
    LP, Lab = sP.l, sa.l + sb.l
    out = np.zeros((len(cart_list(LP)), len(cart_list(sa.l)), len(cart_list(sb.l))))
    sign = _signs(Lab)
    for ip in range(len(sP.exps)):
        g, HP = _aux_hermite(sP, ip)
        for ia in range(len(sa.exps)):
            for ib in range(len(sb.exps)):
                p, Pc, Hab = _pair_hermite(sa, sb, ia, ib)
                alpha = g * p / (g + p)
                R = hermite_R(LP + Lab, alpha, *(sP.center - Pc))
                Rs = _shift_loop(R, LP, Lab)
                f = sP.coefs[ip] * sa.coefs[ia] * sb.coefs[ib] * 2 * np.pi ** 2.5 / (g * p * math.sqrt(g + p))
                out += f * np.einsum("ctuv,xyTUV,tuvTUV->cxy", HP[:, :LP + 1, :LP + 1, :LP + 1], Hab * sign, Rs)
    return _sph(out, (LP, sa.l, sb.l))


def coulomb4_block(sa, sb, sc, sd):
    """(ab|cd) over two orbital shell pairs."""
    L1, L2 = sa.l + sb.l, sc.l + sd.l
    out = np.zeros(tuple(len(cart_list(s.l)) for s in (sa, sb, sc, sd)))
    sign = _signs(L2)
    for ia in range(len(sa.exps)):
        for ib in range(len(sb.exps)):
            p, P, H1 = _pair_hermite(sa, sb, ia, ib)
            for ic in range(len(sc.exps)):
                for id_ in range(len(sd.exps)):
                    q, Q, H2 = _pair_hermite(sc, sd, ic, id_)
                    alpha = p * q / (p + q)
                    R = hermite_R(L1 + L2, alpha, *(P - Q))
                    Rs = _shift_loop(R, L1, L2)
                    f = (sa.coefs[ia] * sb.coefs[ib] * sc.coefs[ic] * sd.coefs[id_] * 2 * np.pi ** 2.5
                         / (p * q * math.sqrt(p + q)))
                    out += f * np.einsum("abtuv,cdTUV,tuvTUV->abcd", H1, H2 * sign, Rs)
    return _sph(out, (sa.l, sb.l, sc.l, sd.l))


# ------------------------------------------------------------ matrices ----

def _norms(shells):
    return np.concatenate([1.0 / np.sqrt(np.diag(overlap_block(s, s))) for s in shells])


def overlap(shells):
    n, off = nbf(shells), offsets(shells)
    S = np.zeros((n, n))
    for i, a in enumerate(shells):
        for j, b in enumerate(shells):
            S[off[i]:off[i] + 2 * a.l + 1, off[j]:off[j] + 2 * b.l + 1] = overlap_block(a, b)
    d = _norms(shells)
    return S * d[:, None] * d[None, :]


def kinetic(shells):
    n, off = nbf(shells), offsets(shells)
    T = np.zeros((n, n))
    for i, a in enumerate(shells):
        for j, b in enumerate(shells):
            T[off[i]:off[i] + 2 * a.l + 1, off[j]:off[j] + 2 * b.l + 1] = kinetic_block(a, b)
    d = _norms(shells)
    return T * d[:, None] * d[None, :]


def dipole(shells, origin=(0.0, 0.0, 0.0)):
    """AO dipole integrals D[d][mu][nu] = <mu| r_d - origin_d |nu>, normalised
    like S (the matrices psi4's MintsHelper.ao_dipole gives the reference's
    Boys localisation, syshfwplug/pymodule.py:156-160, with origin 0)."""
    n, off = nbf(shells), offsets(shells)
    D = np.zeros((3, n, n))
    for i, a in enumerate(shells):
        for j, b in enumerate(shells):
            D[:, off[i]:off[i] + 2 * a.l + 1, off[j]:off[j] + 2 * b.l + 1] = dipole_block(a, b, origin)
    d = _norms(shells)
    return D * d[None, :, None] * d[None, None, :]


def bf_atom(shells) -> np.ndarray:
    """Atom index of every basis function (psi4's function_to_center)."""
    return np.concatenate([np.full(2 * s.l + 1, s.atom, dtype=np.int64) for s in shells])


def nuclear(shells, charges, coords):
    n, off = nbf(shells), offsets(shells)
    V = np.zeros((n, n))
    for i, a in enumerate(shells):
        for j, b in enumerate(shells):
            V[off[i]:off[i] + 2 * a.l + 1, off[j]:off[j] + 2 * b.l + 1] = nuclear_block(a, b, charges, coords)
    d = _norms(shells)
    return V * d[:, None] * d[None, :]


def coulomb2(aux):
    n, off = nbf(aux), offsets(aux)
    M = np.zeros((n, n))
    for i, a in enumerate(aux):
        for j, b in enumerate(aux):
            if j > i:
                continue
            blk = coulomb2_block(a, b)
            M[off[i]:off[i] + 2 * a.l + 1, off[j]:off[j] + 2 * b.l + 1] = blk
            M[off[j]:off[j] + 2 * b.l + 1, off[i]:off[i] + 2 * a.l + 1] = blk.T
    d = _norms(aux)
    return M * d[:, None] * d[None, :]


def coulomb3(aux, shells):
    """(P|mu nu) as [naux, N, N]."""
    na, n = nbf(aux), nbf(shells)
    oa, ob = offsets(aux), offsets(shells)
    T = np.zeros((na, n, n))
    for p, sP in enumerate(aux):
        for i, a in enumerate(shells):
            for j, b in enumerate(shells):
                if j > i:
                    continue
                blk = coulomb3_block(sP, a, b)
                T[oa[p]:oa[p] + 2 * sP.l + 1, ob[i]:ob[i] + 2 * a.l + 1, ob[j]:ob[j] + 2 * b.l + 1] = blk
                T[oa[p]:oa[p] + 2 * sP.l + 1, ob[j]:ob[j] + 2 * b.l + 1, ob[i]:ob[i] + 2 * a.l + 1] = \
                    blk.transpose(0, 2, 1)
    da, d = _norms(aux), _norms(shells)
    return T * da[:, None, None] * d[None, :, None] * d[None, None, :]


def coulomb4(shells):
    """Dense (mu nu|la si), from the unique shell quartets."""
    n, off = nbf(shells), offsets(shells)
    G = np.zeros((n, n, n, n))
    ns = len(shells)
    sl = [slice(off[i], off[i] + 2 * s.l + 1) for i, s in enumerate(shells)]
    for i in range(ns):
        for j in range(i + 1):
            for k in range(ns):
                for l_ in range(k + 1):
                    if (k, l_) > (i, j):
                        continue
                    blk = coulomb4_block(shells[i], shells[j], shells[k], shells[l_])
                    for (a, b, c, d_), perm in (((i, j, k, l_), (0, 1, 2, 3)), ((j, i, k, l_), (1, 0, 2, 3)),
                                                ((i, j, l_, k), (0, 1, 3, 2)), ((j, i, l_, k), (1, 0, 3, 2)),
                                                ((k, l_, i, j), (2, 3, 0, 1)), ((l_, k, i, j), (3, 2, 0, 1)),
                                                ((k, l_, j, i), (2, 3, 1, 0)), ((l_, k, j, i), (3, 2, 1, 0))):
                        G[sl[a], sl[b], sl[c], sl[d_]] = blk.transpose(perm)
    d = _norms(shells)
    return G * d[:, None, None, None] * d[None, :, None, None] * d[None, None, :, None] * d[None, None, None, :]


def nuclear_repulsion(charges, coords) -> float:
    e = 0.0
    for i in range(len(charges)):
        for j in range(i):
            e += charges[i] * charges[j] / np.linalg.norm(coords[i] - coords[j])
    return e
