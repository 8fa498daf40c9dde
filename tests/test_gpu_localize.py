"""AO dipole integrals and orbital localisation on the GPU against the oracle."""
import sys

import numpy as np
import pytest
import torch

from corr_amd.basis import HERE, Molecule, ShellArrays, read_g94, water
from oracle import integrals as I
from oracle import localize as LO

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mol(n):
    """n waters on a line 5.5 bohr apart, each rotated differently."""
    syms, X = [], []
    for k in range(n):
        c, s = np.cos(0.7 * k), np.sin(0.7 * k)
        rot = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
        sy, x = water(rot, (5.5 * k, 0.3 * k, -0.2 * k))
        syms += sy
        X.append(x)
    return Molecule(syms, np.vstack(X))


def _occupied(S, n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((S.shape[0], n))
    w, V = np.linalg.eigh(X.T @ S @ X)
    return X @ V @ np.diag(w ** -0.5) @ V.T


def test_dipole_integrals_vs_oracle():
    import corr_amd.integrals as GI
    mol = _mol(2)
    _say("dipole integrals (H2O)2 def2-SVP")
    O = ShellArrays(mol, read_g94("def2-svp"))
    origin = (0.4, -1.3, 0.9)
    got = GI.dipole(O, origin).cpu().numpy()
    sh = I.build_shells(list(zip(mol.symbols, mol.coords)), I.parse_g94(f"{HERE}/def2-svp.g94"))
    ref = I.dipole(sh, origin)
    assert np.abs(got - ref).max() < 1e-12
    assert np.array_equal(O.bf_atom, I.bf_atom(sh))


def _match(La, Lb, S):
    """max |overlap| of every orbital with its best partner in the other set."""
    M = np.abs(La @ S @ Lb.T)
    return M.max(axis=1), M.max(axis=0)


@pytest.mark.parametrize("kind", ["pm", "boys"])
@pytest.mark.parametrize("nw", [1, 2])
def test_localize_vs_oracle(kind, nw):
    """GPU Jacobi sweeps against the oracle: the same maximum spread as the
    oracle's Jacobi (1e-10 relative) and, for one water (n = 5), as the
    reference's BFGS; the same orbitals up to order and sign; the occupied
    space is unchanged; two calls give identical orbitals."""
    import corr_amd
    import corr_amd.integrals as GI
    _say(f"localize {kind} (H2O){nw}")
    mol = _mol(nw)
    O = ShellArrays(mol, read_g94("def2-svp"))
    sh = I.build_shells(list(zip(mol.symbols, mol.coords)), I.parse_g94(f"{HERE}/def2-svp.g94"))
    S, D, ba = I.overlap(sh), I.dipole(sh), I.bf_atom(sh)
    n = 5 * nw
    C = _occupied(S, n, seed=nw + 7)          # nbf x n (columns)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda")   # noqa: E731
    Crows = dev(C.T)
    if kind == "pm":
        info, L = corr_amd.localize_PM(Crows, dev(S), O.bf_atom)
        spread = lambda X: LO.calculate_pm(S, X, ba)   # noqa: E731
        Q = LO.pm_matrices(S, C, ba, len(mol.symbols))
    else:
        info, L = corr_amd.localize_boys(Crows, GI.dipole(O))
        spread = lambda X: LO.calculate_boys(D, X)     # noqa: E731
        Q = np.stack([C.T @ D[d] @ C for d in range(3)])
    Lg = L.cpu().numpy()
    assert info["converged"]
    assert abs(info["spread_initial"] - spread(C)) < 1e-11 * spread(C)
    assert abs(info["spread"] - spread(Lg.T)) < 1e-11 * info["spread"]
    Lj, _, fj, _ = LO.jacobi_localize(Q, C)
    assert abs(info["spread"] - fj) < 1e-10 * fj
    if nw == 1:
        if kind == "pm":
            _, Lb = LO.localize_bfgs(C, spread, seed=1)
        else:
            _, Lb = LO.localize_bfgs(C, spread, random_start=False)
        assert abs(info["spread"] - spread(Lb)) < 1e-9 * fj
    a, b = _match(Lg, Lj.T, S)
    assert np.allclose(a, 1.0, atol=1e-6) and np.allclose(b, 1.0, atol=1e-6)
    assert np.allclose(Lg @ S @ Lg.T, np.eye(n), atol=1e-12)
    assert np.allclose(Lg.T @ Lg, C @ C.T, atol=1e-11)     # same occupied space
    info2, L2 = corr_amd.localize(Crows, kind, S=dev(S), bf_atom=O.bf_atom, D=GI.dipole(O))
    assert torch.equal(L, L2)


def test_localize_single_orbital_and_errors():
    import corr_amd
    from corr_amd._abi import CorrError
    mol = _mol(1)
    O = ShellArrays(mol, read_g94("def2-svp"))
    sh = I.build_shells(list(zip(mol.symbols, mol.coords)), I.parse_g94(f"{HERE}/def2-svp.g94"))
    S = I.overlap(sh)
    C = _occupied(S, 1, seed=3)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda")   # noqa: E731
    info, L = corr_amd.localize_PM(dev(C.T), dev(S), O.bf_atom)
    assert info["converged"] and torch.equal(L.cpu(), torch.from_numpy(np.ascontiguousarray(C.T)))
    bad = O.bf_atom.copy()
    bad[0] = -1
    with pytest.raises(CorrError):
        corr_amd.localize_PM(dev(C.T), dev(S), bad)
