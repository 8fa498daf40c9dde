"""Pins the Gaussian-integral oracle (oracle/integrals.py) against textbook
values, closed forms, radial quadrature and invariances.  No GPU needed."""
import math
import os

import numpy as np
import pytest
from scipy import integrate

from oracle import integrals as I
from oracle import posthf as O
from oracle import scf as SCF

BAS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "corr_amd", "basis")


def basis(name):
    return I.parse_g94(os.path.join(BAS, name))


def test_sto3g_h2_szabo_ostlund():
    """Szabo & Ostlund, Modern Quantum Chemistry, eqs. 3.229-3.235 (R = 1.4 bohr)."""
    R = 1.4
    X = [np.zeros(3), np.array([0, 0, R])]
    sh = I.build_shells([("H", X[0]), ("H", X[1])], basis("sto-3g.g94"))
    S, T, V, G = I.overlap(sh), I.kinetic(sh), I.nuclear(sh, [1.0, 1.0], X), I.coulomb4(sh)
    assert abs(S[0, 1] - 0.6593) < 5e-5
    assert abs(T[0, 0] - 0.7600) < 5e-5 and abs(T[0, 1] - 0.2365) < 5e-5
    assert abs(V[0, 0] - (-1.2266 - 0.6538)) < 1e-4 and abs(V[0, 1] - 2 * -0.5974) < 1e-4
    for (a, b, c, d), ref in (((0, 0, 0, 0), 0.7746), ((0, 0, 1, 1), 0.5697), ((1, 0, 0, 0), 0.4441),
                              ((1, 0, 1, 0), 0.2970)):
        assert abs(G[a, b, c, d] - ref) < 5e-5
    H = T + V
    r = SCF.run_scf(H, S, lambda P: SCF.fock_dense(G, P, H), SCF.diag_fmat(H, SCF.calc_s12(S))[0], 2,
                    ion_rep=1 / R, e_tol=1e-12)
    assert abs(r["total"] - (-1.1167)) < 5e-5


def test_sto3g_water_published_scf():
    """H2O/STO-3G at the geometry of the Crawford group's programming project
    #3 (bohr): E_nuc = 8.002367061810450, E_SCF = -74.942079928192."""
    X = [np.array([0.0, -0.143225816552, 0.0]), np.array([1.638036840407, 1.136548822547, 0.0]),
         np.array([-1.638036840407, 1.136548822547, 0.0])]
    Z = [8.0, 1.0, 1.0]
    assert abs(I.nuclear_repulsion(Z, X) - 8.002367061810450) < 1e-11
    sh = I.build_shells(list(zip(["O", "H", "H"], X)), basis("sto-3g.g94"))
    S, T, V, G = I.overlap(sh), I.kinetic(sh), I.nuclear(sh, Z, X), I.coulomb4(sh)
    H = T + V
    r = SCF.run_scf(H, S, lambda P: SCF.fock_dense(G, P, H), SCF.diag_fmat(H, SCF.calc_s12(S))[0], 10,
                    ion_rep=I.nuclear_repulsion(Z, X), e_tol=1e-13, maxiter=200)
    assert abs(r["total"] - (-74.942079928192)) < 2e-10


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4])
def test_solid_harmonic_self_overlap_closed_form(l):
    """Raw self-overlap of S_lm e^{-a r^2}: 4 pi/(2l+1) (2l+1)!!/(2^(l+2) (2a)^(l+1)) sqrt(pi/(2a)),
    the same for every m; components of one shell are orthogonal."""
    a = 0.731
    sh = I.Shell(l, np.array([0.2, -0.1, 0.4]), np.array([a]), np.array([1.0]), 0)
    S = I.overlap_block(sh, sh)
    p = 2 * a
    ref = 4 * math.pi / (2 * l + 1) * I._dfact(2 * l + 1) / (2 ** (l + 2) * p ** (l + 1)) * math.sqrt(math.pi / p)
    assert np.allclose(np.diag(S), ref, rtol=1e-13, atol=0)
    assert np.abs(S - np.diag(np.diag(S))).max() < 1e-14 * ref


@pytest.mark.parametrize("l,lp", [(2, 0), (3, 1), (4, 2), (4, 0)])
def test_pure_shells_have_no_lower_l_component(l, lp):
    """Spherical (pure) shells on one centre are orthogonal to every lower l of
    the same parity (a Cartesian d set would overlap the s function)."""
    c = np.array([0.1, 0.3, -0.2])
    a = I.Shell(l, c, np.array([0.9, 0.4]), np.array([0.6, 0.5]), 0)
    b = I.Shell(lp, c, np.array([1.3]), np.array([1.0]), 0)
    assert np.abs(I.overlap_block(a, b)).max() < 1e-15


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_same_centre_coulomb_multipole_quadrature(l):
    """(phi_a | phi_b) for normalised r^l Y_lm e^{-a r^2} on one centre equals
    4 pi/(2l+1) Int Int R_a R_b r_<^l / r_>^(l+1) r1^2 r2^2 (radial quadrature)."""
    a, b = 1.7, 0.45
    c = np.zeros(3)
    sa = I.Shell(l, c, np.array([a]), np.array([1.0]), 0)
    sb = I.Shell(l, c, np.array([b]), np.array([1.0]), 0)
    blk = I.coulomb2_block(sa, sb)
    na = 1 / math.sqrt(I.overlap_block(sa, sa)[0, 0])
    nb = 1 / math.sqrt(I.overlap_block(sb, sb)[0, 0])
    val = blk * na * nb
    Na = 1 / math.sqrt(integrate.quad(lambda r: r ** (2 * l + 2) * math.exp(-2 * a * r * r), 0, np.inf)[0])
    Nb = 1 / math.sqrt(integrate.quad(lambda r: r ** (2 * l + 2) * math.exp(-2 * b * r * r), 0, np.inf)[0])

    def inner(r1):
        f = Na * r1 ** l * math.exp(-a * r1 * r1) * r1 * r1
        lo = integrate.quad(lambda r2: Nb * r2 ** (2 * l + 2) * math.exp(-b * r2 * r2), 0, r1,
                            epsabs=1e-15, epsrel=1e-13)[0] / r1 ** (l + 1)
        hi = integrate.quad(lambda r2: Nb * r2 * math.exp(-b * r2 * r2), r1, np.inf, epsabs=1e-15,
                            epsrel=1e-13)[0] * r1 ** l
        return f * (lo + hi)

    ref = 4 * math.pi / (2 * l + 1) * integrate.quad(inner, 0, np.inf, epsabs=1e-14, epsrel=1e-12, limit=200)[0]
    assert np.allclose(np.diag(val), ref, rtol=1e-9, atol=0)
    assert np.abs(val - np.diag(np.diag(val))).max() < 1e-13


def _water(rot=None, shift=np.zeros(3)):
    r, th = 0.958 * I.ANGSTROM_TO_BOHR, np.deg2rad(104.5)
    X = np.array([[0, 0, 0], [r, 0, 0], [r * np.cos(th), r * np.sin(th), 0]])
    if rot is not None:
        X = X @ rot.T
    X = X + shift
    return list(zip(["O", "H", "H"], X)), [8.0, 1.0, 1.0], list(X)


def _water_energies(rot=None, shift=np.zeros(3)):
    atoms, Z, X = _water(rot, shift)
    sh = I.build_shells(atoms, basis("def2-svp.g94"))
    ax = I.build_shells(atoms, basis("def2-svp-ri.g94"))
    S, T, V, G = I.overlap(sh), I.kinetic(sh), I.nuclear(sh, Z, X), I.coulomb4(sh)
    H = T + V
    r = SCF.run_scf(H, S, lambda P: SCF.fock_dense(G, P, H), SCF.diag_fmat(H, SCF.calc_s12(S))[0], 10,
                    ion_rep=I.nuclear_repulsion(Z, X), e_tol=1e-12, maxiter=200)
    M, _ = O.ri_metric_eigen(I.coulomb2(ax))
    B = O.transform_ri(r["C"], O.ri_bpq(M, I.coulomb3(ax, sh)), 5)
    e = np.diag(O.build_fmo(r["C"], r["F"]))
    return r["total"], O.rimp2(B, e[:5], e[5:])["EMP2"], G


def test_water_def2svp_rotation_translation_invariance():
    e0, m0, G = _water_energies()
    assert np.allclose(G, G.transpose(1, 0, 2, 3)) and np.allclose(G, G.transpose(2, 3, 0, 1))
    q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))
    e1, m1, _ = _water_energies(q, np.array([0.3, -1.1, 2.0]))
    assert abs(e0 - e1) < 1e-10 and abs(m0 - m1) < 1e-11
    assert -76.0 < e0 < -75.9 and -0.21 < m0 < -0.20


def test_ri_exact_for_single_aux_products():
    """When an orbital-basis product is itself one aux function (same centre,
    s x s with exponents summing to the aux exponent), (P|mu nu) = (P|Q)."""
    c = np.array([0.3, 0.2, -0.5])
    s1 = I.Shell(0, c, np.array([0.4]), np.array([1.0]), 0)
    s2 = I.Shell(0, c, np.array([0.7]), np.array([1.0]), 0)
    q = I.Shell(0, c, np.array([1.1]), np.array([1.0]), 0)
    p = I.Shell(2, np.array([1.0, -0.4, 0.9]), np.array([0.8]), np.array([1.0]), 0)
    three = I.coulomb3_block(p, s1, s2)[:, 0, 0]
    two = I.coulomb2_block(p, q)[:, 0]
    assert np.allclose(three, two, rtol=1e-13, atol=1e-16)
