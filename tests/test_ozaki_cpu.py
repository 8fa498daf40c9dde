"""The int8-sliced FP64 GEMM's CPU restatement (the checker of
CORR_PREC_INT8X5) pinned by what the arithmetic fixes: the slices represent
every value to 2^-36 of its row scale, products of small integers come out
exact, random products stay within the slice error bound of the FP64
product, and non-finite entries poison exactly their row or column."""
import numpy as np
import pytest

from tests import ozaki_emulation as OZ


def test_slices_represent_values():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-0.996, 0.996, 4000), rng.uniform(-1, 1, 1000) * 1e-9, [0.0, 0.5, -0.5]])
    q = OZ.slices(x)
    back = sum(qs * 2.0 ** (-7 * (s + 1)) for s, qs in enumerate(q))
    assert np.max(np.abs(back - x)) <= 2.0 ** -36
    assert np.max(np.abs(q[0])) <= 127
    assert all(np.max(np.abs(qs)) <= 64 for qs in q[1:])


def test_scale_headroom():
    mx = np.array([1.0, 0.75, 127.0 / 128, 127.6 / 128, 3e-300, 0.0])
    sc = OZ.scale(mx, np.zeros(6, bool))
    assert np.all(mx / sc * 128 < 127.5)
    assert np.all(np.log2(sc) == np.round(np.log2(sc)))
    assert sc[-1] == 1.0
    assert np.isnan(OZ.scale(np.array([np.finfo(float).max]), np.zeros(1, bool)))[0]


def test_small_integers_exact():
    rng = np.random.default_rng(2)
    A = rng.integers(-127, 128, (37, 300)).astype(float)
    B = rng.integers(-127, 128, (300, 53)).astype(float)
    assert np.array_equal(OZ.gemm_int8x5(A, B), A @ B)


@pytest.mark.parametrize("spread", [0.0, 8.0])
def test_error_bound_against_fp64(spread):
    """|C - AB| <= 4 K 2^-35 sig tau; a dropped diagonal (2^-28) or one slice
    less would break it."""
    rng = np.random.default_rng(3)
    M, K, N = 40, 500, 60
    A = rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-spread, 0, (M, K))
    B = rng.standard_normal((K, N)) * 10.0 ** rng.uniform(-spread, 0, (K, N))
    C = OZ.gemm_int8x5(A, B)
    sig, tau = OZ.row_scales(A), OZ.row_scales(B.T)
    err = np.abs(C - A @ B) / (sig[:, None] * tau[None, :])
    assert np.max(err) <= 4 * K * 2.0 ** -35
    assert np.max(err) > 0


def test_non_finite_rows_and_columns():
    rng = np.random.default_rng(4)
    A, B = rng.standard_normal((6, 20)), rng.standard_normal((20, 5))
    A[2, 7] = np.nan
    B[3, 4] = np.inf
    C = OZ.gemm_int8x5(A, B)
    assert np.all(np.isnan(C[2])) and np.all(np.isnan(C[:, 4]))
    ok = np.ones_like(C, bool)
    ok[2] = False
    ok[:, 4] = False
    assert np.all(np.isfinite(C[ok]))
