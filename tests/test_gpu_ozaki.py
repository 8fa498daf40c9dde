"""The int8-emulated FP64 GEMM (CORR_PREC_INT8X5) and the mixed path's B
formation that uses it (second AO->MO GEMM and Cholesky metric product),
through the C-ABI: bit-identical to the CPU
restatement in tests/ozaki_emulation.py (exact integer slice products, FP64
combination in the kernel's order), within the slice error bound of the FP64
product, NaN rows and columns for non-finite inputs."""
import sys

import numpy as np
import pytest
import torch

from tests import _synth as S
from tests import ozaki_emulation as OZ

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.fixture(scope="module")
def corr():
    import corr_amd
    return corr_amd


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (37, 53, 29), (128, 128, 32), (130, 257, 70), (300, 64, 400),
                                   (257, 384, 1000), (70, 200, 1500), (40, 150, 2500), (33, 129, 3100),
                                   (48, 96, 1024), (48, 96, 1025), (40, 64, 2048), (40, 64, 2049),
                                   (24, 40, 3072), (24, 40, 3073)])
def test_gemm_int8x5_bit_exact(corr, M, N, K):
    """B (k-strided columns) goes through the one-pass slicing kernel held in
    1, 2 or 3 chunk rounds for K <= 1024, 2048, 3072 (each bound and one past
    it tested), and through the two-pass kernel above."""
    _say(f"gemm int8x5 M={M} N={N} K={K}")
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-6, 0, (M, K))
    B = rng.standard_normal((K, N)) * 10.0 ** rng.uniform(-6, 0, (K, N))
    got = corr.dgemm(dev(A), dev(B), alpha=-0.75, prec="int8x5").cpu().numpy()
    want = OZ.gemm_int8x5(A, B, alpha=-0.75)
    assert np.array_equal(got, want), np.max(np.abs(got - want))
    sig, tau = OZ.row_scales(A), OZ.row_scales(B.T)
    err = np.max(np.abs(got + 0.75 * (A @ B)) / (sig[:, None] * tau[None, :]))
    _say(f"  max |C - AB| / (sig tau) = {err:.3e} (bound {4 * K * 2.0 ** -35:.3e})")
    assert err <= 0.75 * 4 * K * 2.0 ** -35


@pytest.mark.parametrize("ta,tb", [(0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(37, 53, 29), (300, 150, 333)])
def test_gemm_int8x5_transposes(corr, ta, tb, M, N, K):
    """Operands read transposed (the other slicing kernels) give the bits of
    the restatement on op(A), op(B)."""
    _say(f"gemm int8x5 ta={ta} tb={tb} M={M} N={N} K={K}")
    rng = np.random.default_rng(M + 3 * N + 5 * K + ta + 2 * tb)
    opA = rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-4, 0, (M, K))
    opB = rng.standard_normal((K, N)) * 10.0 ** rng.uniform(-4, 0, (K, N))
    A = opA.T.copy() if ta else opA
    B = opB.T.copy() if tb else opB
    got = corr.dgemm(dev(A), dev(B), transa=bool(ta), transb=bool(tb), prec="int8x5").cpu().numpy()
    assert np.array_equal(got, OZ.gemm_int8x5(opA, opB))


def test_gemm_int8x5_non_finite(corr):
    rng = np.random.default_rng(11)
    A, B = rng.standard_normal((70, 90)), rng.standard_normal((90, 150))
    A[5, 40] = np.nan
    A[66, 0] = np.inf
    B[17, 129] = -np.inf
    got = corr.dgemm(dev(A), dev(B), prec="int8x5").cpu().numpy()
    bad = np.zeros(got.shape, bool)
    bad[[5, 66]] = True
    bad[:, 129] = True
    assert np.all(np.isnan(got[bad])) and np.all(np.isfinite(got[~bad]))
    assert np.array_equal(got[~bad], OZ.gemm_int8x5(A, B)[~bad])


def test_gemm_int8x5_rejects(corr):
    from corr_amd._abi import CorrError
    a = torch.zeros(2, 75001, dtype=torch.float64, device="cuda")
    b = torch.zeros(75001, 2, dtype=torch.float64, device="cuda")
    with pytest.raises(CorrError):
        corr.dgemm(a, b, prec="int8x5")
    c = torch.ones(3, 3, dtype=torch.float64, device="cuda")
    assert torch.equal(corr.dgemm(c, c, prec="int8x5"), torch.full((3, 3), 3.0, dtype=torch.float64,
                                                                   device="cuda"))


def test_mixed_b_formation_exact_chain(corr):
    """The mixed path's B formation end to end on inputs its int8 steps carry
    exactly: C = I and integer (P|mu nu) with every (P, i) row maximum 100, so
    (P|ia) = (P|o+a, i) comes out of the FP64 and int8 AO->MO GEMMs unrounded;
    the Cholesky metric product (int8, lower-triangular K cut) then gives the
    bits of the restatement on M and that (P|ia)."""
    naux, N, o = 200, 56, 9
    v = N - o
    _say(f"mixed B formation exact chain naux={naux} N={N} o={o}")
    rng = np.random.default_rng(5)
    Y = rng.integers(-60, 61, (naux, N, N)).astype(float)
    Y = np.trunc((Y + Y.transpose(0, 2, 1)) / 2)
    for i in range(N):
        Y[:, i, i] = 100.0
    J = S.metric(naux, seed=6, cond=1e4)
    dJ, dY, dI = dev(J), dev(Y), dev(np.eye(N))
    M, _ = corr.ri_metric(dJ, "cholesky")
    Bm = corr.ri_form_bia(dI, dY, M, o, "cholesky", "bf16x3").cpu().numpy()
    X = Y[:, o:, :o].transpose(0, 2, 1).reshape(naux, o * v)
    assert np.array_equal(Bm.reshape(naux, o * v), OZ.gemm_int8x5(M.cpu().numpy(), X))
    got, _ = corr.ri_bia_from_j(dI, dY, dJ, o, "cholesky", "bf16x3")
    assert np.array_equal(got.cpu().numpy(), Bm)


@pytest.mark.parametrize("naux,N,o", [(152, 48, 10), (300, 70, 13)])
def test_mixed_b_formation(corr, naux, N, o):
    """B formation of the mixed path (prec bf16x3: first AO->MO GEMM in FP64,
    the second and the Cholesky metric product on int8 MFMA) within 1e-7 of
    the FP64 B (relative to its maximum; 4e-10 measured on these well
    conditioned inputs); the overlapped entry gives the same bits."""
    _say(f"mixed B formation naux={naux} N={N} o={o}")
    rng = np.random.default_rng(naux + N)
    J = S.metric(naux, seed=naux, cond=1e5)
    Pmn = rng.standard_normal((naux, N, N))
    Pmn = Pmn + Pmn.transpose(0, 2, 1)
    C = S.orthonormal(N, seed=N)
    dJ, dP, dC = dev(J), dev(Pmn), dev(C)
    M, _ = corr.ri_metric(dJ, "cholesky")
    B64 = corr.ri_form_bia(dC, dP, M, o, "cholesky", "fp64").cpu().numpy()
    Bm = corr.ri_form_bia(dC, dP, M, o, "cholesky", "bf16x3").cpu().numpy()
    rel = np.max(np.abs(Bm - B64)) / np.max(np.abs(B64))
    _say(f"  max |B_mixed - B_fp64| / max |B| = {rel:.3e}")
    assert rel <= 1e-7
    got, _ = corr.ri_bia_from_j(dC, dP, dJ, o, "cholesky", "bf16x3")
    assert np.array_equal(got.cpu().numpy(), Bm)


@pytest.mark.parametrize("naux,N,o,mode,prec", [
    (160, 72, 8, "cholesky", "bf16x3"),     # full panels: o % 4 == 0, v % 32 == 0, naux % 16 == 0
    (150, 61, 7, "cholesky", "bf16x3"),     # ragged o, v and naux
    (300, 90, 13, "cholesky", "bf16x3"),    # several row and column tiles
    (150, 61, 7, "cholesky", "fp64"),
    (150, 61, 7, "cholesky", "fp32"),
    (150, 61, 7, "eigen", "bf16x3"),
])
def test_ri_mp2_energy_from_j(corr, naux, N, o, mode, prec):
    """B formation + energy in one call gives the bits of ri_bia_from_j
    followed by rimp2_energy at the same precision; on the mixed path
    (cholesky, bf16x3) the metric product writes the energy kernel's BF16
    hi/lo panels itself."""
    _say(f"ri_mp2_energy_from_j naux={naux} N={N} o={o} {mode} {prec}")
    v = N - o
    rng = np.random.default_rng(naux * N + o)
    J = S.metric(naux, seed=naux + 1, cond=1e5)
    Pmn = rng.standard_normal((naux, N, N))
    Pmn = Pmn + Pmn.transpose(0, 2, 1)
    C = S.orthonormal(N, seed=N + 1)
    eo = -1.0 - rng.random(o)
    ev = 0.5 + rng.random(v)
    dJ, dP, dC, deo, dev_ = dev(J), dev(Pmn), dev(C), dev(eo), dev(ev)
    Bia, kept = corr.ri_bia_from_j(dC, dP, dJ, o, mode, prec)
    want = corr.rimp2_energy(Bia, deo, dev_, prec)
    got, kept2 = corr.ri_mp2_energy_from_j(dC, dP, dJ, o, deo, dev_, mode, prec)
    assert kept2 == kept
    _say(f"  E = {got['EMP2']:.15f}")
    assert got == want
    # a second call on the same context reuses the panels it zeroed
    again, _ = corr.ri_mp2_energy_from_j(dC, dP, dJ, o, deo, dev_, mode, prec)
    assert again == want


def test_ri_mp2_energy_from_j_rejects(corr):
    dJ = dev(S.metric(40, seed=1, cond=10.0))
    dP, dC = dev(np.zeros((40, 12, 12))), dev(np.eye(12))
    with pytest.raises(ValueError):
        corr.ri_mp2_energy_from_j(dC, dP, dJ, 4, dev(np.zeros(3)), dev(np.ones(8)))
