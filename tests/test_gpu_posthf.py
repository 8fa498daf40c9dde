"""GPU parity of the post-HF kernels against the CPU oracle (through the C-ABI).

Tolerances: FP64 paths agree with the FP64 oracle to 1e-12 relative (energies
to 1e-10 Eh absolute); FP32 MFMA paths to 1e-5 relative on matrix entries;
BF16x3 and FP32 energies to 2e-6 relative on random tensors and to 0.1 uEh
per water on real-chemistry B."""
import sys

import numpy as np
import pytest
import torch

from oracle import posthf as O
from tests import _synth as S

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


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (37, 53, 29), (130, 257, 70), (300, 64, 16), (64, 192, 192)])
def test_gemm_fp64(corr, ta, tb, M, N, K):
    _say(f"gemm fp64 ta={ta} tb={tb} M={M} N={N} K={K}")
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    ref = (A.T if ta else A) @ (B.T if tb else B)
    out = corr.dgemm(dev(A), dev(B), bool(ta), bool(tb)).cpu().numpy()
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * np.sqrt(K))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(200, 150, 300), (37, 53, 29), (130, 257, 70), (300, 64, 33)])
def test_gemm_fp32(corr, ta, tb, M, N, K):
    _say(f"gemm fp32 ta={ta} tb={tb} M={M} N={N} K={K}")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    ref = (A.T if ta else A) @ (B.T if tb else B)
    out = corr.dgemm(dev(A), dev(B), bool(ta), bool(tb), prec="fp32").cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-5 * np.sqrt(K) * 4


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("M,N,K", [(40, 152, 20000), (2000, 1, 3001), (1600, 1300, 2500)])
def test_gemm_split_k(corr, ta, tb, prec, M, N, K):
    """Fewer output tiles than CUs and K >= 2048 take the split-K path (chunks
    of K into a workspace, fixed-order sum; the last chunk is ragged): a thin
    deep product, a matrix-vector product and a medium output."""
    _say(f"gemm split-K {prec} ta={ta} tb={tb} M={M} N={N} K={K}")
    rng = np.random.default_rng(11 + 2 * ta + tb)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    ref = -0.5 * ((A.T if ta else A) @ (B.T if tb else B))
    out = corr.dgemm(dev(A), dev(B), bool(ta), bool(tb), alpha=-0.5, prec=prec).cpu().numpy()
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert np.max(np.abs(out - ref)) < tol * np.sqrt(K) * 4
    again = corr.dgemm(dev(A), dev(B), bool(ta), bool(tb), alpha=-0.5, prec=prec).cpu().numpy()
    assert np.array_equal(out, again)   # deterministic


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gemm_large_output_tiles(corr, ta, tb, prec):
    """An output of >= 512 tiles of 128 x 128 takes the weighted tile choice
    (128 x 192, 12 waves); ragged in both M and N."""
    M, N, K = 3001, 2999, 77
    _say(f"gemm large output {prec} ta={ta} tb={tb}")
    rng = np.random.default_rng(21 + ta)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    ref = (A.T if ta else A) @ (B.T if tb else B)
    out = corr.dgemm(dev(A), dev(B), bool(ta), bool(tb), prec=prec).cpu().numpy()
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert np.max(np.abs(out - ref)) < tol * np.sqrt(K) * 4


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_fp64_staged_interior_tiles(corr, ta, tb):
    """Even leading dimensions and 16-byte aligned operands: the interior
    tiles stage by global_load_lds, the K remainder (1014 = 63 x 16 + 6) and
    the edge tiles through registers; batched with even strides.  The same
    product from operands one double off 16-byte alignment takes the register
    path only; both match numpy."""
    from corr_amd import _abi
    from corr_amd._abi import Context, ptr
    M, N, K, nb = 520, 400, 1014, 3
    _say(f"gemm fp64 staged ta={ta} tb={tb} M={M} N={N} K={K} batch={nb}")
    rng = np.random.default_rng(31 + 2 * ta + tb)
    ash, bsh = ((K, M) if ta else (M, K)), ((N, K) if tb else (K, N))
    A = rng.standard_normal((nb,) + ash)
    B = rng.standard_normal((nb,) + bsh)
    ref = np.stack([(A[b].T if ta else A[b]) @ (B[b].T if tb else B[b]) for b in range(nb)])
    ctx = Context.get(0)
    for off in (0, 1):
        dA = torch.zeros(A.size + 2, dtype=torch.float64, device="cuda")
        dB = torch.zeros(B.size + 2, dtype=torch.float64, device="cuda")
        dA[off:off + A.size] = dev(A.ravel())
        dB[off:off + B.size] = dev(B.ravel())
        C = torch.zeros(nb, M, N, dtype=torch.float64, device="cuda")
        ctx.call("corr_dgemm", ta, tb, M, N, K, 1.0, ptr(dA[off:]), ash[1], A[0].size, ptr(dB[off:]), bsh[1],
                 B[0].size, 0.0, ptr(C), N, M * N, nb, _abi.PRECISIONS["fp64"])
        got = C.cpu().numpy()
        assert np.max(np.abs(got - ref)) < 1e-12 * np.sqrt(K) * 4, off


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("shape", ["128x160 of 2x5", "256x128 of 4x4", "128x192 of 2x6"])
def test_gemm_fp64_wide_wave_tiles(corr, ta, tb, shape):
    """Outputs of >= 512 tiles whose padded areas select the FP64 tiles with
    2 x 5 (N = 150: 128 x 160), 4 x 4 (N = 7552 = 59 x 128: 256 x 128) and
    2 x 6 (N = 170, one column of 128 x 192) MFMA blocks per wave; ragged M, a K remainder below one slab (K = 2 x 16 + 10),
    batched; 16-byte aligned operands with even leading dimensions (staged
    interior tiles) and operands one double off alignment (register path),
    both against numpy."""
    from corr_amd import _abi
    from corr_amd._abi import Context, ptr
    M, N, K, nb = {"128x160 of 2x5": (66554, 150, 42, 1), "256x128 of 4x4": (2554, 7552, 42, 2),
                   "128x192 of 2x6": (66554, 170, 42, 1)}[shape]
    _say(f"gemm fp64 {shape} ta={ta} tb={tb} M={M} N={N} K={K} batch={nb}")
    rng = np.random.default_rng(41 + 2 * ta + tb)
    ash, bsh = ((K, M) if ta else (M, K)), ((N, K) if tb else (K, N))
    A = rng.standard_normal((nb,) + ash)
    B = rng.standard_normal((nb,) + bsh)
    ref = np.stack([(A[b].T if ta else A[b]) @ (B[b].T if tb else B[b]) for b in range(nb)])
    ctx = Context.get(0)
    for off in (0, 1):
        dA = torch.zeros(A.size + 2, dtype=torch.float64, device="cuda")
        dB = torch.zeros(B.size + 2, dtype=torch.float64, device="cuda")
        dA[off:off + A.size] = dev(A.ravel())
        dB[off:off + B.size] = dev(B.ravel())
        C = torch.zeros(nb, M, N, dtype=torch.float64, device="cuda")
        ctx.call("corr_dgemm", ta, tb, M, N, K, 1.0, ptr(dA[off:]), ash[1], A[0].size, ptr(dB[off:]), bsh[1],
                 B[0].size, 0.0, ptr(C), N, M * N, nb, _abi.PRECISIONS["fp64"])
        err = float(np.max(np.abs(C.cpu().numpy() - ref)))
        _say(f"  offset {off}: max |diff| = {err:.2e}")
        assert err < 1e-12 * np.sqrt(K) * 4, off


@pytest.mark.parametrize("entry", ["ri_metric", "ri_bia_from_j"])
def test_cholesky_metric_not_positive_definite(corr, entry):
    """A metric with a negative eigenvalue is refused with CORR_ENOTPD (the
    factorisation flags a non-positive pivot); the context keeps working, also
    after the failure happened on the second stream of ri_bia_from_j."""
    from corr_amd._abi import CORR_ENOTPD, CorrError
    naux, N, o = 64, 16, 4
    _say(f"cholesky not-PD via {entry}")
    rng = np.random.default_rng(91)
    q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    lam = np.geomspace(1.0, 1e3, naux)
    lam[naux // 2] = -0.5
    J = (q * lam) @ q.T
    J = 0.5 * (J + J.T)
    with pytest.raises(CorrError) as err:
        if entry == "ri_metric":
            corr.ri_metric(dev(J), "cholesky")
        else:
            Pmn = rng.standard_normal((naux, N, N))
            corr.ri_bia_from_j(dev(S.orthonormal(N, seed=92)), dev(Pmn + Pmn.transpose(0, 2, 1)), dev(J), o,
                               "cholesky", "fp64")
    assert err.value.status == CORR_ENOTPD
    B = S.bia(40, 3, 9, seed=93)
    eo, ev = S.orbital_energies(3, 9, seed=93)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), "fp64")
    assert abs(got["EMP2"] - O.rimp2(B, eo, ev)["EMP2"]) < 1e-12


def test_calls_on_two_streams_are_ordered(corr):
    """One context serves every torch stream of a device and its scratch is
    shared, while ri_form_bia returns before its kernels finish.  Calls issued
    back to back on two streams (no host synchronisation between them) must
    give the same bits as the same calls run one after the other."""
    naux, N, o = 608, 192, 40
    _say("ri_form_bia on two torch streams")
    rng = np.random.default_rng(81)
    ins = []
    for seed in (82, 83):
        J = S.metric(naux, seed=seed, cond=1e4)
        P = rng.standard_normal((naux, N, N))
        ins.append((dev(S.orthonormal(N, seed=seed)), dev(P + P.transpose(0, 2, 1)), dev(J)))
    serial = []
    for C, P, J in ins:
        M, _ = corr.ri_metric(J, "cholesky")
        serial.append((M, corr.ri_form_bia(C, P, M, o, "cholesky", "fp32").cpu()))
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for s in (s1, s2):
        s.wait_stream(torch.cuda.current_stream())
    outs = []
    for rep in range(3):
        for (C, P, _), (M, _), s in zip(ins, serial, (s1, s2)):
            with torch.cuda.stream(s):
                outs.append(corr.ri_form_bia(C, P, M, o, "cholesky", "fp32"))
    torch.cuda.synchronize()
    for n, B in enumerate(outs):
        assert torch.equal(B.cpu(), serial[n % 2][1]), n


@pytest.mark.parametrize("mode", ["cholesky", "eigen"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ri_bia_from_j_equals_two_calls(corr, mode, prec):
    """The overlapped entry (metric on the second stream) gives exactly the
    bits of ri_metric followed by ri_form_bia."""
    naux, N, o = 152, 48, 10
    _say(f"ri_bia_from_j {mode} {prec}")
    rng = np.random.default_rng(77)
    J = S.metric(naux, seed=78, cond=1e5)
    Pmn = rng.standard_normal((naux, N, N))
    Pmn = Pmn + Pmn.transpose(0, 2, 1)
    C = S.orthonormal(N, seed=79)
    dJ, dP, dC = dev(J), dev(Pmn), dev(C)
    M, kept = corr.ri_metric(dJ, mode)
    ref = corr.ri_form_bia(dC, dP, M, o, mode, prec).cpu().numpy()
    got, kept2 = corr.ri_bia_from_j(dC, dP, dJ, o, mode, prec)
    assert kept2 == kept
    assert np.array_equal(got.cpu().numpy(), ref)


@pytest.mark.parametrize("naux,o,v", [(76, 5, 19), (50, 1, 1), (33, 9, 40), (140, 13, 70), (304, 20, 76)])
def test_rimp2_fp64_matches_oracle(corr, naux, o, v):
    _say(f"rimp2 fp64 naux={naux} o={o} v={v}")
    B = S.bia(naux, o, v, seed=naux + o)
    eo, ev = S.orbital_energies(o, v, seed=o + v)
    ref = O.rimp2(B, eo, ev)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec="fp64")
    for k in ref:
        assert abs(got[k] - ref[k]) < 1e-10 and abs(got[k] - ref[k]) <= 1e-12 * abs(ref[k]) + 1e-15, k


def test_rimp2_many_tiles_two_pass_reduction(corr):
    """o = 128, v = 256: 528 x 36 = 19008 workgroup partials, past the 16384
    that one reduction workgroup takes, so the partials go through the fixed
    256-chunk first pass.  FP64 against the oracle, and repeatable."""
    naux, o, v = 64, 128, 256
    nib, nac = (o + 3) // 4, (v + 31) // 32
    assert nib * (nib + 1) // 2 * (nac * (nac + 1) // 2) > 16384
    _say(f"rimp2 fp64 two-pass reduction naux={naux} o={o} v={v}")
    B = S.bia(naux, o, v, seed=11)
    eo, ev = S.orbital_energies(o, v, seed=12)
    ref = O.rimp2(B, eo, ev)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec="fp64")
    for k in ref:
        assert abs(got[k] - ref[k]) <= 1e-12 * abs(ref[k]) + 1e-15, k
    assert corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec="fp64") == got


@pytest.mark.parametrize("prec,rtol", [("fp32", 2e-6), ("bf16x3", 2e-6)])
def test_rimp2_reduced_precision(corr, prec, rtol):
    naux, o, v = 304, 20, 76
    _say(f"rimp2 {prec}")
    B = S.bia(naux, o, v, seed=5)
    eo, ev = S.orbital_energies(o, v, seed=6)
    ref = O.rimp2(B, eo, ev)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec=prec)
    for k in ref:
        assert abs(got[k] - ref[k]) <= rtol * abs(ref[k]), (k, got[k], ref[k])


@pytest.mark.parametrize("naux,o,v", [(20, 13, 31), (33, 5, 64), (48, 9, 150), (40, 16, 224), (64, 21, 608)])
def test_rimp2_bf16x3_exact_products(corr, naux, o, v):
    """BF16x3 tile coverage and weights, sharply: B holds small integers
    (exact in BF16, every M^{ij}_ab sum exact in FP32), so the BF16x3 energy
    equals the FP64 oracle's to rounding of the FP64 epilogue.  The virtual
    block counts 1, 2, 5, 7 and 19 give a diagonal block alone, one pair of
    diagonal blocks, pairs with an odd block left over (in a group of four and
    in a group of three), and the (H2O)32 layout; odd occupied block counts
    leave the last 8-block half empty."""
    _say(f"rimp2 bf16x3 exact products naux={naux} o={o} v={v}")
    rng = np.random.default_rng(naux * o + v)
    B = rng.integers(-8, 9, (naux, o, v)).astype(float)
    eo, ev = S.orbital_energies(o, v, seed=o + v)
    ref = O.rimp2(B, eo, ev)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec="bf16x3")
    for k in ref:
        assert abs(got[k] - ref[k]) <= 1e-12 * abs(ref[k]), (k, got[k], ref[k])


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_rimp2_deterministic(corr, prec):
    """Fixed-order reductions: repeated calls give bit-identical energies in
    every precision."""
    naux, o, v = 140, 13, 70
    B, (eo, ev) = S.bia(naux, o, v, seed=9), S.orbital_energies(o, v, seed=9)
    a = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec)
    for _ in range(3):
        assert corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec) == a


def _water_fixture():
    from pathlib import Path
    w = np.load(Path(__file__).resolve().parent / "golden" / "rimp2_water_def2svp.npz")
    return w["Bia"], w["eps_occ"], w["eps_vir"], {"EMP2_OS": float(w["e_os"]), "EMP2_SS": float(w["e_ss"]),
                                                  "EMP2": float(w["e_mp2"]), "EMP2_SCS": float(w["e_scs"])}


def _separated_waters(n, seed):
    """n copies of the water's B with no overlap between molecules (B^Q_ia = 0
    unless Q, i and a belong to one water), occupied, virtual and auxiliary
    indices each shuffled: M^{ij}_ab vanishes across molecules, so EMP2 is
    exactly n times the water's (size extensivity).  Realistic magnitudes at
    o = 5n, v = 19n, naux = 76n."""
    Bw, eo, ev, e = _water_fixture()
    naux1, o1, v1 = Bw.shape
    B = np.zeros((n * naux1, n * o1, n * v1))
    for m in range(n):
        B[m * naux1:(m + 1) * naux1, m * o1:(m + 1) * o1, m * v1:(m + 1) * v1] = Bw
    rng = np.random.default_rng(seed)
    po, pv, pq = rng.permutation(n * o1), rng.permutation(n * v1), rng.permutation(n * naux1)
    B = B[pq][:, po][:, :, pv]
    return B, np.tile(eo, n)[po], np.tile(ev, n)[pv], {k: n * x for k, x in e.items()}


@pytest.mark.parametrize("prec,bound_uEh", [("fp64", 1e-4), ("fp32", 0.05), ("bf16x3", 0.3)])
def test_rimp2_water_def2svp_micro_hartree(corr, prec, bound_uEh):
    """Real one-water B (def2-SVP / def2-SVP-RI, oracle integrals and RHF,
    eigen metric): every energy component within an absolute bound in
    micro-hartree (measured: FP32 0.004, BF16x3 0.125)."""
    B, eo, ev, ref = _water_fixture()
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec=prec)
    err = max(abs(got[k] - ref[k]) for k in ref) * 1e6
    _say(f"rimp2 water def2-SVP {prec}: max |diff| = {err:.6f} uEh")
    assert err <= bound_uEh


@pytest.mark.parametrize("prec,bound_uEh", [("fp64", 1e-4), ("fp32", 0.4), ("bf16x3", 2.4)])
def test_rimp2_separated_waters_size_extensive(corr, prec, bound_uEh):
    """Eight separated waters (o = 40, v = 152, naux = 608, shuffled indices):
    EMP2 = 8 x EMP2(water) exactly; the reduced precisions within the
    single-water bounds times 8."""
    B, eo, ev, ref = _separated_waters(8, seed=3)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec=prec)
    err = max(abs(got[k] - ref[k]) for k in ref) * 1e6
    _say(f"rimp2 8 separated waters {prec}: max |diff| = {err:.6f} uEh")
    assert err <= bound_uEh


@pytest.mark.parametrize("naux", [7, 64, 65, 200, 700])
def test_ri_metric(corr, naux):
    """Both factors against the oracle; at 700 the Cholesky trailing updates
    skip tiles above the diagonal and the inverse's L21 X11 products start
    their K range at the tile's first column, over several 128-row tiles."""
    _say(f"ri_metric naux={naux}")
    J = S.metric(naux, seed=naux, cond=1e5)
    Me, kept = corr.ri_metric(dev(J), "eigen")
    Mr, _ = O.ri_metric_eigen(J)
    assert kept == naux
    assert np.allclose(Me.cpu().numpy(), Mr, rtol=1e-9, atol=1e-9)
    Mc, _ = corr.ri_metric(dev(J), "cholesky")
    assert np.allclose(Mc.cpu().numpy(), O.ri_metric_cholesky(J), rtol=1e-9, atol=1e-9)


def test_ri_metric_cutoff(corr):
    naux = 40
    q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((naux, naux)))
    lam = np.geomspace(1.0, 10.0, naux)
    lam[:3] = [1e-14, -1e-13, 5e-12]
    J = (q * lam) @ q.T
    M, kept = corr.ri_metric(dev(J), "eigen")
    assert kept == naux - 3
    Mr, _ = O.ri_metric_eigen(J)
    assert np.allclose(M.cpu().numpy(), Mr, atol=1e-8)


def test_ri_metric_not_pd_raises(corr):
    J = -np.eye(10)
    with pytest.raises(corr.CorrError):
        corr.ri_metric(dev(J), "cholesky")


@pytest.mark.parametrize("naux,N,o", [(20, 9, 3), (76, 24, 5), (130, 40, 7)])
def test_ri_b_formation(corr, naux, N, o):
    _say(f"ri B naux={naux} N={N} o={o}")
    rng = np.random.default_rng(N)
    C = S.orthonormal(N, seed=N)
    X = rng.standard_normal((naux, N, N)) * 0.1
    Pmn = 0.5 * (X + X.transpose(0, 2, 1))
    J = S.metric(naux, seed=o, cond=1e4)
    Me, _ = O.ri_metric_eigen(J)
    BPQ = O.ri_bpq(Me, Pmn)
    ref = O.transform_ri(C, BPQ, o)
    got_bpq = corr.calculate_ri(dev(J), dev(Pmn), "eigen")
    assert np.allclose(got_bpq.cpu().numpy(), BPQ, rtol=1e-9, atol=1e-10)
    got = corr.transform_ri(dev(C), got_bpq, o)
    assert np.allclose(got.cpu().numpy(), ref, rtol=1e-9, atol=1e-10)
    # reordered formation (transform first, metric last), both metric kinds
    Mc = O.ri_metric_cholesky(J)
    for M, mode in [(Me, "eigen"), (Mc, "cholesky")]:
        Bf = corr.ri_form_bia(dev(C), dev(Pmn), dev(M), o, mode=mode).cpu().numpy()
        ref_f = np.einsum("PQ,Qia->Pia", M, O.transform_ri(C, Pmn, o))
        assert np.allclose(Bf, ref_f, rtol=1e-10, atol=1e-11)


def test_build_fmo(corr):
    N = 37
    rng = np.random.default_rng(2)
    C, F = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    got = corr.build_FMo(dev(C), dev(F)).cpu().numpy()
    assert np.allclose(got, O.build_fmo(C, F), rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("N,o", [(6, 2), (13, 4), (24, 5)])
def test_unfold_motrans_canonical_mp2(corr, N, o):
    _say(f"motrans N={N}")
    dense = S.eri_like(N, seed=N)
    val, idx, sc = O.sorted_unique_list(dense)
    d_idx = torch.from_numpy(idx.view(np.int16).copy()).to("cuda")
    ao = corr.ao_unfold(N, dev(val), d_idx, sc)
    assert torch.equal(ao.cpu(), torch.from_numpy(dense))
    C = S.orthonormal(N, seed=N + 1)
    ref = O.motrans(C, dense)
    mo = corr.MOtrans(dev(C), ao)
    assert np.allclose(mo.cpu().numpy(), ref, rtol=1e-12, atol=1e-13)
    eo, ev = S.orbital_energies(o, N - o, seed=N)
    FMo = np.diag(np.concatenate([eo, ev]))
    e_ref = O.canonical_mp2(ref, FMo, o)
    e = corr.run_canonical_mp2(mo, dev(FMo), o)
    for k in e_ref:
        assert abs(e[k] - e_ref[k]) < 1e-12, k
    mo32 = corr.MOtrans(dev(C), ao, prec="fp32").cpu().numpy()
    assert np.max(np.abs(mo32 - ref)) < 1e-5 * np.max(np.abs(ref))


def test_unfold_shuffled_screened_list(corr):
    """A list in random order, with a fifth of the entries screened out and a
    quarter stored with the pairs swapped ((cd|ab) for (ab|cd)): every stored
    value lands on its 8 images and the images of absent entries are zero
    (P = 861 pairs, not a multiple of the 64-wide mirror tiles)."""
    N = 41
    _say(f"unfold shuffled/screened N={N}")
    dense = S.eri_like(N, seed=5)
    val, idx, sc = O.sorted_unique_list(dense)
    rng = np.random.default_rng(6)
    keep = rng.random(len(val)) > 0.2
    val, idx = val[keep], idx[keep].astype(np.int64)
    swap = rng.random(len(val)) < 0.25
    idx[swap] = idx[swap][:, [2, 3, 0, 1]]
    order = rng.permutation(len(val))
    val, idx = val[order], idx[order]
    ref = np.zeros_like(dense)
    a, b, c, d = idx.T
    for p_, q_, r_, s_ in [(a, b, c, d), (b, a, c, d), (a, b, d, c), (b, a, d, c),
                           (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a)]:
        ref[p_, q_, r_, s_] = val
    d_idx = torch.from_numpy(idx.astype(np.uint16).view(np.int16).copy()).to("cuda")
    sc_all = np.array([0, 0, 0, 0], dtype=np.int64)
    ao = corr.ao_unfold(N, dev(val), d_idx, sc_all)
    assert torch.equal(ao.cpu(), torch.from_numpy(ref))


@pytest.mark.parametrize("N,o,rotate", [(10, 3, False), (10, 3, True), (37, 9, True)])
def test_non_canonical_mp2_vs_oracle(N, o, rotate):
    """run_non_canonical_mp2 (loc_mp2.cc:7-92): 20 iterations of E(sem-loc)
    and Rsum against the oracle (1e-12 relative); with rotated occupied
    orbitals (exp of a random antisymmetric generator, angle 0.3) the energy
    converges to the canonical MP2 energy of the converged RHF."""
    print(f"non-canonical MP2 N={N} o={o} rotate={rotate}", flush=True)
    from scipy.linalg import expm
    import corr_amd
    *_, h, eri, _X = S.nonorthogonal_system(N, seed=N, nocc=o)
    C, eps, F = S.rhf_orthonormal(h, eri, o)
    e_can = O.canonical_mp2(O.motrans(C, eri), O.build_fmo(C, F), o)["EMP2"]
    if rotate:
        A = np.random.default_rng(N + 1).standard_normal((o, o))
        C = C.copy()
        C[:o] = expm(0.3 * (A - A.T)) @ C[:o]
    prec = O.motrans(C, eri)
    FMo = O.build_fmo(C, F)
    E, Rs, T = O.non_canonical_mp2(prec, FMo, o, 20)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda")   # noqa: E731
    r = corr_amd.run_non_canonical_mp2(d(prec), d(FMo), o, 20, return_T=True)
    assert np.allclose(r["energies"], E, rtol=1e-12, atol=1e-14)
    assert np.allclose(r["rsum"], Rs, rtol=1e-9, atol=1e-13)
    assert np.allclose(r["T"].cpu().numpy(), T, rtol=0, atol=1e-13)
    assert abs(r["energies"][-1] - e_can) < 1e-10
