"""GPU RI-MP2 gradient: against the oracle's analytic FP64 gradient, against
central finite differences of the GPU FP64 energy (h = 1e-4, seeded
entries), and through torch.autograd.  FP64 tolerance 1e-10 relative to the
largest gradient entry; FP32 MFMA pair blocks 1e-5."""
import sys

import numpy as np
import pytest
import torch

from oracle import posthf as O
from tests import _synth as S

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.fixture(scope="module")
def corr():
    import corr_amd
    return corr_amd


@pytest.mark.parametrize("naux,o,v,c_os,c_ss", [(40, 5, 19, 1.0, 1.0), (76, 9, 45, 1.2, 1.0 / 3.0),
                                                (60, 3, 70, 1.0, 1.0)])
def test_gradient_matches_oracle(corr, naux, o, v, c_os, c_ss):
    print(f"grad naux={naux} o={o} v={v}", flush=True)
    B = S.bia(naux, o, v, seed=o + v)
    eo, ev = S.orbital_energies(o, v, seed=naux)
    E, dB, deo, dev_ = O.rimp2_gradient(B, eo, ev, c_os, c_ss)
    gE, gB, geo, gev = corr.rimp2_gradient(dev(B), dev(eo), dev(ev), c_os, c_ss, prec="fp64")
    assert abs(gE - E) < 1e-12 * abs(E) + 1e-14
    sc = np.abs(dB).max()
    assert np.abs(gB.cpu().numpy() - dB).max() < 1e-10 * sc
    assert np.abs(geo.cpu().numpy() - deo).max() < 1e-10 * np.abs(deo).max()
    assert np.abs(gev.cpu().numpy() - dev_).max() < 1e-10 * np.abs(dev_).max()
    fE, fB, feo, fev = corr.rimp2_gradient(dev(B), dev(eo), dev(ev), c_os, c_ss, prec="fp32")
    assert np.abs(fB.cpu().numpy() - dB).max() < 1e-5 * sc


def test_gradient_central_differences(corr):
    naux, o, v = 64, 6, 40
    B = S.bia(naux, o, v, seed=3)
    eo, ev = S.orbital_energies(o, v, seed=4)
    _, gB, geo, gev = corr.rimp2_gradient(dev(B), dev(eo), dev(ev), prec="fp64")
    gB = gB.cpu().numpy()
    h = 1e-4
    rng = np.random.default_rng(64)
    d_eo, d_ev = dev(eo), dev(ev)
    for _ in range(64):
        q, i, a = rng.integers(naux), rng.integers(o), rng.integers(v)
        Bp, Bm = B.copy(), B.copy()
        Bp[q, i, a] += h
        Bm[q, i, a] -= h
        ep = corr.rimp2_energy(dev(Bp), d_eo, d_ev)["EMP2"]
        em = corr.rimp2_energy(dev(Bm), d_eo, d_ev)["EMP2"]
        fd = (ep - em) / (2 * h)
        assert abs(gB[q, i, a] - fd) < 1e-7 * max(1.0, abs(fd)), (q, i, a, gB[q, i, a], fd)


def test_autograd_function(corr):
    naux, o, v = 50, 4, 30
    B = dev(S.bia(naux, o, v, seed=8)).requires_grad_(True)
    eo_, ev_ = S.orbital_energies(o, v, seed=8)
    eo = dev(eo_).requires_grad_(True)
    ev = dev(ev_).requires_grad_(True)
    E = corr.rimp2(B, eo, ev, prec="fp64")
    E.backward()
    E_ref, dB, deo, dev_ = O.rimp2_gradient(B.detach().cpu().numpy(), eo_, ev_)
    assert abs(E.item() - E_ref) < 1e-12 * abs(E_ref)
    assert np.abs(B.grad.cpu().numpy() - dB).max() < 1e-10 * np.abs(dB).max()
    assert np.abs(eo.grad.cpu().numpy() - deo).max() < 1e-10 * np.abs(deo).max()
    assert np.abs(ev.grad.cpu().numpy() - dev_).max() < 1e-10 * np.abs(dev_).max()


def test_fused_and_blocked_paths_agree(corr):
    """Up to half the context's scratch limit R (o v x o v) is formed whole by
    the pair-tile kernel; above it the occupied-blocked path runs.  Both give
    the oracle's gradient; reruns of the fused path are bit-identical."""
    from corr_amd._abi import Context, lib
    naux, o, v = 120, 13, 100     # ragged: 4 occupied blocks, 4 virtual blocks
    print(f"fused vs blocked naux={naux} o={o} v={v}", flush=True)
    B = S.bia(naux, o, v, seed=5)
    eo, ev = S.orbital_energies(o, v, seed=6)
    E, dB, deo, dev_ = O.rimp2_gradient(B, eo, ev, 1.2, 1.0 / 3.0)
    d = (dev(B), dev(eo), dev(ev))
    fused = corr.rimp2_gradient(*d, 1.2, 1.0 / 3.0, prec="fp64")
    again = corr.rimp2_gradient(*d, 1.2, 1.0 / 3.0, prec="fp64")
    for x, y in zip(fused[1:], again[1:]):
        assert torch.equal(x, y)
    assert fused[0] == again[0]
    ctx = Context.get(0)
    ov = o * v
    ctx.check("corr_ctx_trim", lib().corr_ctx_trim(ctx.handle))
    ctx.check("corr_ctx_set_scratch_limit", lib().corr_ctx_set_scratch_limit(ctx.handle, 8 * ov * ov))
    try:
        blocked = corr.rimp2_gradient(*d, 1.2, 1.0 / 3.0, prec="fp64")
    finally:
        ctx.check("corr_ctx_set_scratch_limit", lib().corr_ctx_set_scratch_limit(ctx.handle, 0))
    sc = np.abs(dB).max()
    for g in (fused, blocked):
        assert abs(g[0] - E) < 1e-12 * abs(E)
        assert np.abs(g[1].cpu().numpy() - dB).max() < 1e-10 * sc
        assert np.abs(g[2].cpu().numpy() - deo).max() < 1e-10 * np.abs(deo).max()
        assert np.abs(g[3].cpu().numpy() - dev_).max() < 1e-10 * np.abs(dev_).max()
    f32 = corr.rimp2_gradient(*d, 1.2, 1.0 / 3.0, prec="fp32")
    assert abs(f32[0] - E) < 1e-6 * abs(E)
    assert np.abs(f32[1].cpu().numpy() - dB).max() < 1e-5 * sc
    assert np.abs(f32[3].cpu().numpy() - dev_).max() < 1e-5 * np.abs(dev_).max()
