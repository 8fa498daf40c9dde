"""Basis sets and molecules for the GPU integral engine.

Reads Gaussian-94 basis files (the format of the reference's basissets/*.g94;
the H and O blocks of STO-3G, def2-SVP and def2-SVP-RI ship in corr_amd/basis)
and lays a basis on a molecule as the plain shell arrays of corr_basis
(include/corr_abi.h): spherical shells, atom order, then file order.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
ANGSTROM = 1.0 / 0.52917721067          # bohr per angstrom (CODATA 2014)
Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
_LCHAR = "SPDFGHI"


def read_g94(name_or_path: str) -> dict:
    """{element: [(l, exponents, coefficients)]}; an SP shell becomes s + p."""
    path = name_or_path if os.path.exists(name_or_path) else os.path.join(HERE, name_or_path.lower() + ".g94")
    out, el = {}, None
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        i += 1
        if not tok or tok[0].startswith("!"):
            continue
        if tok[0] == "****":
            el = None
            continue
        if el is None:
            if tok[0] in Z and len(tok) > 1 and tok[1] == "0":
                el = tok[0]
                out[el] = []
            continue
        kind, n = tok[0].upper(), int(tok[1])
        rows = np.array([[float(x.replace("D", "E")) for x in lines[i + k].split()] for k in range(n)])
        i += n
        if kind == "SP":
            out[el] += [(0, rows[:, 0], rows[:, 1]), (1, rows[:, 0], rows[:, 2])]
        else:
            out[el].append((_LCHAR.index(kind), rows[:, 0], rows[:, 1]))
    return out


@dataclass
class Molecule:
    symbols: list
    coords: np.ndarray        # (natom, 3) bohr

    @property
    def charges(self) -> np.ndarray:
        return np.array([float(Z[s]) for s in self.symbols])

    def nuclear_repulsion(self) -> float:
        q, X = self.charges, self.coords
        e = 0.0
        for a in range(len(q)):
            for b in range(a):
                e += q[a] * q[b] / float(np.linalg.norm(X[a] - X[b]))
        return e

    def nelectrons(self) -> int:
        return int(self.charges.sum())


class ShellArrays:
    """A basis laid on a molecule, kept alive for the corr_basis pointers."""

    def __init__(self, mol: Molecule, basis: dict):
        l, npr, cen, ex, co = [], [], [], [], []
        for sym, xyz in zip(mol.symbols, mol.coords):
            for (ll, e, c) in basis[sym]:
                l.append(ll)
                npr.append(len(e))
                cen.extend(np.asarray(xyz, float))
                ex.extend(e)
                co.extend(c)
        atom = []
        for a, sym in enumerate(mol.symbols):
            for (ll, e, c) in basis[sym]:
                atom.extend([a] * (2 * ll + 1))
        self.bf_atom = np.array(atom, dtype=np.int32)   # atom of every basis function
        self.l = np.array(l, dtype=np.int32)
        self.nprim = np.array(npr, dtype=np.int32)
        self.center = np.array(cen, dtype=np.float64)
        self.exps = np.array(ex, dtype=np.float64)
        self.coefs = np.array(co, dtype=np.float64)
        self.nbf = int(np.sum(2 * self.l + 1))

    def struct(self):
        from ._abi import CorrBasis
        p = lambda a, t: a.ctypes.data_as(C.POINTER(t))   # noqa: E731
        return CorrBasis(len(self.l), p(self.l, C.c_int32), p(self.nprim, C.c_int32), p(self.center, C.c_double),
                         p(self.exps, C.c_double), p(self.coefs, C.c_double))


def water(rot: np.ndarray | None = None, shift=(0.0, 0.0, 0.0)):
    """H2O with O-H 0.958 A and H-O-H 104.5 deg: (symbols, coords in bohr)."""
    r, th = 0.958 * ANGSTROM, np.deg2rad(104.5)
    X = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0], [r * np.cos(th), r * np.sin(th), 0.0]])
    if rot is not None:
        X = X @ np.asarray(rot).T
    return ["O", "H", "H"], X + np.asarray(shift, float)


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(n: int, seed: int = 0, volume_per_molecule_A3: float = 40.0, min_oo_A: float = 2.8) -> Molecule:
    """n rigid waters: O centres uniform in a cube of edge (n V)^(1/3),
    rejecting O-O < min_oo, each molecule rotated by a uniform random
    quaternion; seeded."""
    rng = np.random.default_rng(seed)
    L = (n * volume_per_molecule_A3) ** (1.0 / 3.0) * ANGSTROM
    dmin = min_oo_A * ANGSTROM
    centres = []
    while len(centres) < n:
        c = rng.uniform(0.0, L, 3)
        if all(np.linalg.norm(c - d) >= dmin for d in centres):
            centres.append(c)
    syms, X = [], []
    for c in centres:
        s, x = water(_random_rotation(rng), c)
        syms += s
        X.append(x)
    return Molecule(syms, np.vstack(X))
