"""corr_rimp2_energy_noalloc (the allocation-free RI-MP2 energy, workspace in
a static device array) through the C-ABI: on two calls in a row, the second
with a smaller problem, it returns the bits of corr_rimp2_energy in every
precision; a problem whose workspace exceeds the static array is rejected
with CORR_EINVAL and a message naming the size."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from tests import _synth as S

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def call(entry, B, eo, ev, prec):
    from corr_amd import _abi
    ctx = _abi.Context.get(B.device.index or 0)
    e = _abi.MP2Energy()
    naux, o, v = B.shape
    ctx.bind_stream()
    st = getattr(_abi.lib(), entry)(ctx.handle, naux, o, v, _abi.ptr(B), _abi.ptr(eo), _abi.ptr(ev),
                                    _abi.PRECISIONS[prec], C.byref(e))
    msg = _abi.lib().corr_ctx_last_error(ctx.handle)
    return st, e.as_dict(), (msg.decode() if msg else "")


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_noalloc_matches_energy_on_two_calls(prec):
    from corr_amd import _abi
    for naux, o, v, seed in ((608, 40, 152, 1), (200, 13, 50, 2)):
        print(f"noalloc {prec} naux={naux} o={o} v={v}", flush=True)
        sys.stdout.flush()
        B = dev(S.bia(naux, o, v, seed=seed))
        eo_h, ev_h = S.orbital_energies(o, v, seed=seed)
        eo, ev = dev(eo_h), dev(ev_h)
        st, got, msg = call("corr_rimp2_energy_noalloc", B, eo, ev, prec)
        assert st == _abi.CORR_OK, msg
        st, ref, msg = call("corr_rimp2_energy", B, eo, ev, prec)
        assert st == _abi.CORR_OK, msg
        assert got == ref, (got, ref)


def test_noalloc_rejects_workspace_beyond_the_static_array():
    from corr_amd import _abi
    naux, o, v = 2432, 160, 608   # (H2O)32: ~1 GB of panels
    B = torch.zeros(naux, o, v, dtype=torch.float64, device="cuda")
    eo = torch.full((o,), -1.0, dtype=torch.float64, device="cuda")
    ev = torch.full((v,), 1.0, dtype=torch.float64, device="cuda")
    st, _, msg = call("corr_rimp2_energy_noalloc", B, eo, ev, "bf16x3")
    print("rejected:", msg, flush=True)
    assert st == _abi.CORR_EINVAL and "static scratch holds" in msg
    # the context stays usable
    st, _, msg = call("corr_rimp2_energy_noalloc", B[:64, :8, :32].contiguous(), eo[:8].contiguous(),
                      ev[:32].contiguous(), "fp64")
    assert st == _abi.CORR_OK, msg
