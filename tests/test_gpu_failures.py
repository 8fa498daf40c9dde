"""Failure paths on the GPU.  The reference has none: its IO and libint paths
call exit() (ops_io.cc:288-289, ops_rhf.cc:251) and nothing else is checked.
Here every entry point returns a status and the context stays usable:

  NaN in B or in the orbital energies propagates: the energies come back
  NaN (no status; the caller checks isfinite) in every precision.  An
  infinite orbital energy removes its terms (1/inf = 0), as in the reference.
  NaN in the metric J: CORR_ENOTPD from the Cholesky path (a non-finite
  pivot), CORR_ENOTPD (non-finite eigenvalues) or CORR_ENOTCONV (dsyevd) from
  the eigen path.
  A workspace beyond the context's scratch limit, or a device allocation the
  GPU cannot give: CORR_ENOMEM, and the next call on the context succeeds."""
import ctypes as C
import math
import sys

import numpy as np
import pytest
import torch

from oracle import posthf as O
from tests import _synth as S

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.fixture(scope="module")
def corr():
    import corr_amd
    return corr_amd


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("where", ["B", "eps_occ", "eps_vir"])
def test_nan_inputs_give_nan_energies(corr, prec, where):
    naux, o, v = 140, 13, 70
    say(f"NaN in {where}, {prec}")
    B = S.bia(naux, o, v, seed=21)
    eo, ev = S.orbital_energies(o, v, seed=21)
    {"B": lambda: B.__setitem__((17, 5, 33), np.nan), "eps_occ": lambda: eo.__setitem__(7, np.nan),
     "eps_vir": lambda: ev.__setitem__(40, np.nan)}[where]()
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec)
    assert all(math.isnan(x) for x in got.values()), got
    # the context is unharmed: a clean call gives the oracle's energies
    B2 = S.bia(naux, o, v, seed=22)
    eo2, ev2 = S.orbital_energies(o, v, seed=22)
    got2 = corr.rimp2_energy(dev(B2), dev(eo2), dev(ev2), "fp64")
    assert abs(got2["EMP2"] - O.rimp2(B2, eo2, ev2)["EMP2"]) < 1e-10


@pytest.mark.parametrize("prec,rtol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-6)])
def test_infinite_orbital_energy_drops_its_terms(corr, prec, rtol):
    """eps_a = +inf makes every denominator with a infinite: IEEE 1/inf = 0
    removes those terms, as in the reference's X = 1/D (can_mp2.cc:95-101),
    in every precision; the oracle with the same input agrees."""
    naux, o, v = 140, 13, 70
    B = S.bia(naux, o, v, seed=23)
    eo, ev = S.orbital_energies(o, v, seed=23)
    ev[40] = np.inf
    ref = O.rimp2(B, eo, ev)
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), prec)
    say(f"inf eps_vir {prec}: {got['EMP2']:.12f} oracle {ref['EMP2']:.12f}")
    assert all(math.isfinite(x) for x in got.values())
    for k in ref:
        assert abs(got[k] - ref[k]) <= rtol * abs(ref[k])


@pytest.mark.parametrize("mode", ["cholesky", "eigen"])
def test_nan_metric_is_reported(corr, mode):
    from corr_amd._abi import CORR_ENOTCONV, CORR_ENOTPD, CorrError
    naux = 96
    say(f"NaN in J, {mode}")
    J = S.metric(naux, seed=31, cond=1e4)
    J[40, 12] = J[12, 40] = np.nan
    with pytest.raises(CorrError) as err:
        corr.ri_metric(dev(J), mode)
    say(f"  {err.value}")
    assert err.value.status in ((CORR_ENOTPD,) if mode == "cholesky" else (CORR_ENOTPD, CORR_ENOTCONV))
    J2 = S.metric(naux, seed=32, cond=1e4)
    M, kept = corr.ri_metric(dev(J2), mode)
    assert kept == naux
    Mh = M.cpu().numpy()
    assert np.allclose(Mh.T @ Mh @ J2, np.eye(naux), atol=1e-8)


def test_scratch_limit_gives_enomem_and_context_recovers(corr):
    from corr_amd import _abi
    from corr_amd._abi import CORR_ENOMEM, Context, CorrError, lib
    naux, o, v = 304, 20, 76
    say("scratch limit")
    B = S.bia(naux, o, v, seed=41)
    eo, ev = S.orbital_energies(o, v, seed=41)
    ctx = Context.get(0)
    ctx.check("corr_ctx_trim", lib().corr_ctx_trim(ctx.handle))
    ctx.check("corr_ctx_set_scratch_limit", lib().corr_ctx_set_scratch_limit(ctx.handle, 1 << 16))
    try:
        with pytest.raises(CorrError) as err:
            corr.rimp2_energy(dev(B), dev(eo), dev(ev), "bf16x3")
        say(f"  {err.value}")
        assert err.value.status == CORR_ENOMEM and "scratch limit" in str(err.value)
    finally:
        ctx.check("corr_ctx_set_scratch_limit", lib().corr_ctx_set_scratch_limit(ctx.handle, 0))
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), "fp64")
    assert abs(got["EMP2"] - O.rimp2(B, eo, ev)["EMP2"]) < 1e-10
    # a device allocation no GPU can give
    p = C.c_void_p()
    st = lib().corr_dmalloc(ctx.handle, C.byref(p), 1 << 52)
    msg = lib().corr_ctx_last_error(ctx.handle).decode()
    say(f"  corr_dmalloc(2^52): status {st}, {msg}")
    assert st == CORR_ENOMEM and not p.value
    got = corr.rimp2_energy(dev(B), dev(eo), dev(ev), "bf16x3")
    assert abs(got["EMP2"] - O.rimp2(B, eo, ev)["EMP2"]) < 2e-6 * abs(got["EMP2"])
    assert _abi.CORR_OK == 0
