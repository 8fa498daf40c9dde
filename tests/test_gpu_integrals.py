"""GPU Gaussian integrals (corr_amd/integrals.py, integrals.hip) against the
integral oracle (oracle/integrals.py).  Bar: 1e-12 absolute (integrals are
O(1)), SCF energy of H2O/STO-3G to the published -74.942079928192 Eh."""
import sys

import numpy as np
import pytest
import torch

from corr_amd import basis as B
from oracle import integrals as I
from oracle import scf as SCF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gi():
    from corr_amd import integrals
    return integrals


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def oracle_shells(mol, name):
    return I.build_shells(list(zip(mol.symbols, mol.coords)), I.parse_g94(f"{B.HERE}/{name}.g94"))


def mono(rot_seed=None):
    rot = None if rot_seed is None else B._random_rotation(np.random.default_rng(rot_seed))
    s, X = B.water(rot, (0.4, -0.3, 1.2))
    return B.Molecule(s, X)


def test_sto3g_water_scf_published(gi):
    say("integrals sto-3g water")
    import corr_amd
    X = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                  [-1.638036840407, 1.136548822547, 0.0]])
    mol = B.Molecule(["O", "H", "H"], X)
    obs = B.ShellArrays(mol, B.read_g94("sto-3g"))
    S, T, V = gi.one_electron(mol, obs)
    G = gi.four_center(obs)
    sh = oracle_shells(mol, "sto-3g")
    assert np.abs(S.cpu().numpy() - I.overlap(sh)).max() < 1e-13
    assert np.abs(T.cpu().numpy() - I.kinetic(sh)).max() < 1e-12
    assert np.abs(V.cpu().numpy() - I.nuclear(sh, mol.charges, list(mol.coords))).max() < 1e-12
    assert np.abs(G.cpu().numpy() - I.coulomb4(sh)).max() < 1e-13
    H = (T + V).contiguous()
    r = corr_amd.run_scf(H, S, G, torch.zeros_like(S), 10, ion_rep=mol.nuclear_repulsion(), e_tol=1e-13,
                         maxiter=200)
    assert abs(r["e_total"] - (-74.942079928192)) < 2e-10


@pytest.mark.parametrize("rot_seed", [None, 3])
def test_water_def2svp_all_integrals(gi, rot_seed):
    say(f"integrals def2-svp water rot={rot_seed}")
    mol = mono(rot_seed)
    obs = B.ShellArrays(mol, B.read_g94("def2-svp"))
    aux = B.ShellArrays(mol, B.read_g94("def2-svp-ri"))
    sh, ax = oracle_shells(mol, "def2-svp"), oracle_shells(mol, "def2-svp-ri")
    S, T, V = gi.one_electron(mol, obs)
    assert np.abs(S.cpu().numpy() - I.overlap(sh)).max() < 1e-13
    assert np.abs(T.cpu().numpy() - I.kinetic(sh)).max() < 1e-12
    assert np.abs(V.cpu().numpy() - I.nuclear(sh, mol.charges, list(mol.coords))).max() < 1e-12
    J = gi.metric(aux).cpu().numpy()
    assert np.abs(J - I.coulomb2(ax)).max() < 1e-12
    P3 = gi.three_center(aux, obs).cpu().numpy()
    assert np.abs(P3 - I.coulomb3(ax, sh)).max() < 1e-12
    G = gi.four_center(obs).cpu().numpy()
    assert np.abs(G - I.coulomb4(sh)).max() < 1e-12


def test_water_dimer_ri_integrals(gi):
    say("integrals def2-svp water dimer")
    mol = B.water_cluster(2, seed=5)
    obs = B.ShellArrays(mol, B.read_g94("def2-svp"))
    aux = B.ShellArrays(mol, B.read_g94("def2-svp-ri"))
    sh, ax = oracle_shells(mol, "def2-svp"), oracle_shells(mol, "def2-svp-ri")
    assert np.abs(gi.metric(aux).cpu().numpy() - I.coulomb2(ax)).max() < 1e-12
    assert np.abs(gi.three_center(aux, obs).cpu().numpy() - I.coulomb3(ax, sh)).max() < 1e-12
    S, T, V = gi.one_electron(mol, obs)
    assert np.abs(V.cpu().numpy() - I.nuclear(sh, mol.charges, list(mol.coords))).max() < 1e-12


def test_h2o32_three_center_spot_checks(gi):
    """Full (H2O)32 shape (N = 768, naux = 2432): 60 random shell triples of
    the GPU tensor against the oracle's block, and the mu <-> nu symmetry."""
    say("integrals (H2O)32 3c spot checks")
    mol = B.water_cluster(32, seed=0)
    obs = B.ShellArrays(mol, B.read_g94("def2-svp"))
    aux = B.ShellArrays(mol, B.read_g94("def2-svp-ri"))
    P3 = gi.three_center(aux, obs)
    assert P3.shape == (2432, 768, 768)
    sh, ax = oracle_shells(mol, "def2-svp"), oracle_shells(mol, "def2-svp-ri")
    o_sh, o_ax = I.offsets(sh), I.offsets(ax)
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(60):
        p, a, b = rng.integers(len(ax)), rng.integers(len(sh)), rng.integers(len(sh))
        ref = I.coulomb3_block(ax[p], sh[a], sh[b])
        na = lambda s: 1 / np.sqrt(np.diag(I.overlap_block(s, s)))   # noqa: E731
        ref = ref * na(ax[p])[:, None, None] * na(sh[a])[None, :, None] * na(sh[b])[None, None, :]
        got = P3[o_ax[p]:o_ax[p] + 2 * ax[p].l + 1, o_sh[a]:o_sh[a] + 2 * sh[a].l + 1,
                 o_sh[b]:o_sh[b] + 2 * sh[b].l + 1].cpu().numpy()
        worst = max(worst, np.abs(got - ref).max())
    assert worst < 1e-12, worst
    idx = torch.randint(0, 2432, (16,), device=P3.device)
    blk = P3[idx]
    assert torch.equal(blk, blk.transpose(1, 2))
    del P3
    torch.cuda.empty_cache()


def test_g_aux_shells_against_oracle(gi):
    """Aux shells up to g (l = 4, as in JK-fitting sets): (P|Q) and (P|mu nu)
    for every (LX, LY) class with LX = 4, on a water with def2-SVP."""
    say("integrals g aux shells")
    mol = mono(5)
    aux = {"O": [(4, np.array([1.9]), np.array([1.0])), (3, np.array([0.8, 2.6]), np.array([0.4, 0.7])),
                 (0, np.array([3.1]), np.array([1.0]))],
           "H": [(4, np.array([1.1]), np.array([1.0])), (2, np.array([0.9]), np.array([1.0]))]}
    obs = B.ShellArrays(mol, B.read_g94("def2-svp"))
    A = B.ShellArrays(mol, aux)
    sh = oracle_shells(mol, "def2-svp")
    ax = I.build_shells(list(zip(mol.symbols, mol.coords)), aux)
    assert np.abs(gi.metric(A).cpu().numpy() - I.coulomb2(ax)).max() < 1e-12
    assert np.abs(gi.three_center(A, obs).cpu().numpy() - I.coulomb3(ax, sh)).max() < 1e-12
