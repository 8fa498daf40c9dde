"""Pins of the localisation oracle (oracle/localize.py) and of the AO dipole
integrals (oracle/integrals.py dipole), CPU only."""
import numpy as np
import pytest

from corr_amd.basis import HERE, water
from oracle import integrals as I
from oracle import localize as LO


def _water_shells(shift=(0.0, 0.0, 0.0)):
    sym, X = water(shift=shift)
    return I.build_shells(list(zip(sym, X)), I.parse_g94(f"{HERE}/def2-svp.g94"))


def _occupied(S, n, seed):
    """n orbitals orthonormal in the S metric (columns), from a random start."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((S.shape[0], n))
    w, V = np.linalg.eigh(X.T @ S @ X)
    return X @ V @ np.diag(w ** -0.5) @ V.T


def _quad_dipole_cart(sa, sb, origin, d, ca, cb):
    """<cart a| r_d - origin_d |cart b> by 1-D trapezoid quadrature per axis."""
    x = np.linspace(-14.0, 14.0, 40001)
    tot = 0.0
    for ea, wa in zip(sa.exps, sa.coefs):
        for eb, wb in zip(sb.exps, sb.coefs):
            prod = wa * wb
            for k in range(3):
                f = (x - sa.center[k]) ** ca[k] * (x - sb.center[k]) ** cb[k] * np.exp(
                    -ea * (x - sa.center[k]) ** 2 - eb * (x - sb.center[k]) ** 2)
                if k == d:
                    f = f * (x - origin[k])
                prod *= np.trapezoid(f, x)
            tot += prod
    return tot


def test_dipole_against_quadrature():
    """A contracted d shell against a p shell on another centre: the Hermite
    form E_1 + X_PC E_0 against direct 1-D quadrature of the Cartesian
    products, through the same solid-harmonic transform."""
    A, B, O = np.array([0.3, -0.2, 0.5]), np.array([-0.4, 0.6, -0.1]), np.array([0.7, 0.1, -0.9])
    sa = I.Shell(2, A, np.array([1.3, 0.4]), np.array([0.8, 0.5]), 0)
    sb = I.Shell(1, B, np.array([0.9, 0.25]), np.array([0.6, 0.7]), 1)
    got = I.dipole_block(sa, sb, O)
    for d in range(3):
        cart = np.array([[_quad_dipole_cart(sa, sb, O, d, ca, cb) for cb in I.cart_list(1)] for ca in I.cart_list(2)])
        ref = I._sph(cart, (2, 1))
        assert np.allclose(got[d], ref, rtol=0, atol=1e-11), d


def test_dipole_translation_and_symmetry():
    """D(origin + t) = D(origin) - t S; moving the molecule by t adds t S;
    the matrices are symmetric."""
    sh = _water_shells()
    S, D0 = I.overlap(sh), I.dipole(sh)
    t = np.array([0.3, -0.7, 1.1])
    Dt = I.dipole(sh, t)
    D2 = I.dipole(_water_shells(shift=t))
    for d in range(3):
        assert np.allclose(Dt[d], D0[d] - t[d] * S, atol=1e-13)
        assert np.allclose(D2[d], D0[d] + t[d] * S, atol=1e-13)
        assert np.allclose(D0[d], D0[d].T, atol=1e-14)


def test_pm_populations_and_invariances():
    sh = _water_shells()
    S, ba = I.overlap(sh), I.bf_atom(sh)
    L = _occupied(S, 5, seed=3)
    P = LO.populations(S, L, ba)
    assert np.allclose(P.sum(axis=0), 1.0, atol=1e-12)      # each orbital's populations sum to 1
    f = LO.calculate_pm(S, L, ba)
    perm = L[:, [3, 0, 4, 2, 1]] * np.array([1, -1, 1, -1, -1])
    assert abs(LO.calculate_pm(S, perm, ba) - f) < 1e-12
    Q = LO.pm_matrices(S, L, ba, 3)
    assert abs(LO.spread_of(Q) - f) < 1e-12                  # Q_A diagonals are the populations


def test_jacobi_angle_is_the_scan_maximum():
    """The closed-form 2x2 angle against a brute-force scan of g."""
    rng = np.random.default_rng(5)
    Q = rng.standard_normal((4, 6, 6))
    Q = Q + Q.transpose(0, 2, 1)
    L = np.eye(6)
    for s, t in [(0, 1), (2, 5), (3, 4)]:
        A, B = LO.pair_terms(Q, s, t)
        g = 0.25 * np.arctan2(B, -A)
        best = LO.spread_of(LO.rotate(Q, L, s, t, g)[0])
        scan = max(LO.spread_of(LO.rotate(Q, L, s, t, x)[0]) for x in np.linspace(-np.pi, np.pi, 4001))
        assert best >= scan - 1e-9
        assert abs(best - LO.spread_of(Q) - (A + np.hypot(A, B))) < 1e-10   # the predicted gain


@pytest.mark.parametrize("kind", ["pm", "boys"])
def test_jacobi_and_reference_bfgs_reach_the_same_maximum(kind):
    """On H2O (n = 5, the size the reference's sizing supports) the reference's
    BFGS over expm(X) and the Jacobi sweeps reach the same spread, with the
    same orbitals up to order and sign; at the maximum every pair term B is 0;
    Boys' trace is rotation invariant."""
    sh = _water_shells()
    S, ba, D = I.overlap(sh), I.bf_atom(sh), I.dipole(sh)
    C = _occupied(S, 5, seed=11)
    if kind == "pm":
        spread = lambda L: LO.calculate_pm(S, L, ba)   # noqa: E731
        Q = LO.pm_matrices(S, C, ba, 3)
        res, Lb = LO.localize_bfgs(C, spread, seed=1)
    else:
        spread = lambda L: LO.calculate_boys(D, L)     # noqa: E731
        Q = np.stack([C.T @ D[d] @ C for d in range(3)])
        res, Lb = LO.localize_bfgs(C, spread, random_start=False)
    Lj, Qj, fj, _ = LO.jacobi_localize(Q, C)
    assert abs(fj - spread(Lj)) < 1e-12
    assert abs(fj - spread(Lb)) < 1e-9 * fj
    M = np.abs(Lj.T @ S @ Lb)
    assert np.allclose(np.sort(M.max(axis=1)), 1.0, atol=1e-4)
    assert np.allclose(np.sort(M.max(axis=0)), 1.0, atol=1e-4)
    for s in range(5):
        for t in range(s + 1, 5):
            assert abs(LO.pair_terms(Qj, s, t)[1]) < 1e-8
    if kind == "boys":
        for d in range(3):
            assert abs(np.trace(Lj.T @ D[d] @ Lj) - np.trace(C.T @ D[d] @ C)) < 1e-12
